"""Pins of the oracle (oracle/oracle.c) against mathematics, CPU only.

Each oracle function is checked against something that does not restate its
own formula: closed forms of one source-target pair, finite differences of the
potential, numpy.fft, a dense kernel-matrix M2L on the actual surfaces, and an
FP64 direct sum for the whole evaluation.
"""
import numpy as np
import pytest

import oracle as O
from fmm_cases import equiv_rel_diff, kernel_np, make_problem, rel_err

KERNELS = [(O.LAPLACE, 0.0), (O.YUKAWA, 3.0), (O.HELMHOLTZ, 4.0 + 1.5j)]


def closed_form(kid, wk, q, s, t):
    d = t - s
    r = np.linalg.norm(d)
    if kid == O.LAPLACE:
        phi = q / (4 * np.pi * r)
        grad = -q * d / (4 * np.pi * r ** 3)
    elif kid == O.YUKAWA:
        k = np.real(wk)
        phi = q * np.exp(-k * r) / (4 * np.pi * r)
        grad = -q * np.exp(-k * r) * (1 + k * r) * d / (4 * np.pi * r ** 3)
    else:
        phi = q * np.exp(1j * wk * r) / (4 * np.pi * r)
        grad = phi * (1j * wk - 1 / r) * d / r
    return phi, grad


@pytest.mark.parametrize("kid,wk", KERNELS)
def test_direct_two_body_closed_form(kid, wk):
    rng = np.random.default_rng(1)
    for _ in range(5):
        s, t = rng.random(3), rng.random(3) + 0.3
        q = 0.7 if kid != O.HELMHOLTZ else 0.7 - 0.2j
        v = O.direct(kid, s[None], np.array([q]), t[None], wk)[0]
        phi, grad = closed_form(kid, wk, q, s, t)
        assert abs(v[0] - phi) <= 1e-14 * abs(phi)
        assert np.linalg.norm(v[1:] - grad) <= 1e-14 * np.linalg.norm(grad)


@pytest.mark.parametrize("kid,wk", KERNELS)
def test_direct_gradient_is_derivative_of_potential(kid, wk):
    # the gradient output is d(phi)/d(target position) (laplace.h:158-161)
    rng = np.random.default_rng(2)
    src = rng.random((40, 3))
    q = rng.random(40) - 0.5
    if kid == O.HELMHOLTZ:
        q = q + 1j * (rng.random(40) - 0.5)
    trg = rng.random((6, 3)) + 1.5
    g = O.direct(kid, src, q, trg, wk)
    h = 1e-5
    for d in range(3):
        e = np.zeros(3)
        e[d] = h
        fd = (O.direct(kid, src, q, trg + e, wk, False) - O.direct(kid, src, q, trg - e, wk, False)) / (2 * h)
        assert rel_err(g[:, 1 + d], fd) < 1e-8


@pytest.mark.parametrize("kid,wk", KERNELS)
def test_direct_coincident_points_and_linearity(kid, wk):
    rng = np.random.default_rng(3)
    src = rng.random((50, 3))
    q1, q2 = rng.random(50), rng.random(50) - 0.5
    trg = np.vstack([src[:5], rng.random((5, 3))])
    a = O.direct(kid, src, q1 + 3 * q2, trg, wk)
    b = O.direct(kid, src, q1, trg, wk) + 3 * O.direct(kid, src, q2, trg, wk)
    assert rel_err(a, b) < 1e-14
    # a target on a source: that pair is skipped (laplace.h:57) -> finite values
    assert np.all(np.isfinite(a))
    only = O.direct(kid, src[:1], q1[:1], src[:1], wk)
    assert np.all(only == 0)


@pytest.mark.parametrize("n", [8, 12, 16, 20])
def test_dft_matches_numpy(n):
    rng = np.random.default_rng(n)
    x = rng.random((n, n, n))
    X = O.dft3_r2c(x)
    assert rel_err(X, np.fft.rfftn(x)) < 1e-13
    # unnormalised backward transform (FFTW c2r convention)
    assert rel_err(O.dft3_c2r(X, n), x * n ** 3) < 1e-13
    z = rng.random((n, n, n)) + 1j * rng.random((n, n, n))
    assert rel_err(O.dft3_c2c(z, -1), np.fft.fftn(z)) < 1e-13
    assert rel_err(O.dft3_c2c(z, +1), np.fft.ifftn(z) * n ** 3) < 1e-13


@pytest.mark.parametrize("p", [2, 3, 4, 6, 8, 10, 12, 16])
def test_surface_and_grid_maps(p):
    n = 6 * (p - 1) ** 2 + 2
    s = O.surface(p, 1.0, 0, np.zeros(3), 1.0)
    assert s.shape == (n, 3)
    # every point on the unit cube surface, all distinct, on the p x p lattice
    assert np.allclose(np.max(np.abs(s), axis=1), 1.0)
    assert len({tuple(np.round(v, 9)) for v in s}) == n
    lat = (s + 1) * (p - 1) / 2
    assert np.allclose(lat, np.round(lat), atol=1e-9)
    # second half is the point reflection of the first (geometry.h:52-54)
    assert np.allclose(s[n // 2:], -s[: n // 2])
    # level / alpha / centre scaling
    s2 = O.surface(p, 2.0, 3, np.array([1.0, 2.0, 3.0]), 2.95)
    assert np.allclose(s2, s * 2.0 * 0.125 * 2.95 + [1, 2, 3])
    n1 = 2 * p
    for dn in (0, 1):
        m = O.surf2conv(p, dn)
        assert len(set(m.tolist())) == n
        ix, iy, iz = m % n1, (m // n1) % n1, m // (n1 * n1)
        lo = p if dn else 0
        for a in (ix, iy, iz):
            assert a.min() >= lo and a.max() <= lo + p - 1
        # index = offset - lattice coordinate (geometry.h:108-110, 129-131)
        off = 2 * p - 1 if dn else p - 1
        assert np.array_equal(np.stack([ix, iy, iz], 1), off - np.round(lat).astype(int))


def _full_tree(kernel):
    """A uniform depth-2 tree: 64 level-2 leaves of 8 points (ncrit 8), all 8
    level-1 nodes non-leaf.  Points sit at +-0.05 around the 4x4x4 cell centres
    so the bounding box and every split are unambiguous."""
    cells = np.array([[i, j, k] for k in range(4) for j in range(4) for i in range(4)], dtype=float)
    corners = np.array([[a, b, c] for c in (-1, 1) for b in (-1, 1) for a in (-1, 1)], dtype=float)
    x = ((cells[:, None, :] + 0.5) / 4.0 + 0.05 * corners[None, :, :]).reshape(-1, 3)
    return x


@pytest.mark.parametrize("kid,wk,p", [k + (4,) for k in KERNELS] + [(O.LAPLACE, 0.0, 8)])
def test_fft_m2l_equals_dense_m2l(kid, wk, p):
    """oracle M2L (operator build + fft_up_equiv + hadamard + ifft_dn_check)
    equals the dense kernel matrix from every non-adjacent source child's
    upward-equivalent surface to the target child's downward-check surface
    (p = 4 for the three kernels, and p = 8, n1 = 16, for Laplace)."""
    name = {O.LAPLACE: "laplace", O.YUKAWA: "modified_helmholtz", O.HELMHOLTZ: "helmholtz"}[kid]
    x = _full_tree(name)
    import exafmm
    from fmm_cases import KERNELS as KC
    mod, cls, _, _ = KC[name]
    Cls = getattr(mod, cls)
    q = np.ones(len(x)) if kid != O.HELMHOLTZ else np.ones(len(x)) + 0j
    fmm = Cls(p, 8, "") if kid == O.LAPLACE else Cls(p, 8, wk, "")
    fmm.filename = ""
    tree = mod.setup(mod.init_sources(x, q), mod.init_targets(x), fmm)
    f = fmm.flat_tree(tree)
    assert f["depth"] == 2 and f["nnodes"] == 73
    r0 = f["r0"]
    lap = kid == O.LAPLACE
    slots = 1 if lap else f["depth"]
    m2l = np.concatenate([O.m2l_operator(kid, p, r0, 0 if lap else l + 1, lap, wk) for l in range(slots)])
    ops = fmm.operators()
    ops["m2l"] = m2l
    orc = O.Oracle(kid, p, f, ops, wk)
    rng = np.random.default_rng(5)
    ns = orc.nsurf
    up = rng.random((f["nnodes"], ns)) - 0.5
    if kid == O.HELMHOLTZ:
        up = up + 1j * (rng.random((f["nnodes"], ns)) - 0.5)
    orc.up[:] = up
    orc.run(O.ST_M2L)
    # dense M2L between level-2 children of M2L pairs of level-1 nodes
    c0, cen, lev = f["child0"], f["center"].reshape(-1, 3), f["level"]
    dense = np.zeros_like(orc.dn)
    for T in np.nonzero(lev == 1)[0]:
        for r in range(26):
            S = f["m2l_source"][T * 26 + r]
            if S < 0:
                continue
            for a in range(8):
                tc = c0[T] + a
                chk = O.surface(p, r0, 2, cen[tc], 1.05)
                for b in range(8):
                    sc = c0[S] + b
                    rel = np.round((cen[sc] - cen[tc]) / (2 * r0 / 4)).astype(int)
                    if np.max(np.abs(rel)) <= 1:
                        continue  # adjacent children: near field, not M2L
                    eq = O.surface(p, r0, 2, cen[sc], 1.05)
                    K = kernel_np(kid, wk, chk[:, None, :] - eq[None, :, :])
                    dense[tc] += K @ up[sc]
    assert rel_err(orc.dn, dense) < 1e-12


@pytest.mark.parametrize("kernel,p,tol", [("laplace", 4, 5e-4), ("laplace", 6, 2e-5), ("modified_helmholtz", 5, 1e-4),
                                          ("helmholtz", 5, 1e-4), ("laplace", 8, 1e-7), ("laplace", 10, 5e-8),
                                          ("modified_helmholtz", 8, 2e-7), ("helmholtz", 8, 2e-7)])
def test_oracle_fmm_against_direct(kernel, p, tol):
    """The oracle's full evaluation against its direct sum (3,000 points, ncrit 32; measured at p = 8 / 10:
    Laplace phi 2.8e-8 / 2.4e-8, grad 1.8e-7 / 1.5e-7; Yukawa and Helmholtz p = 8 phi 5.1e-8 / 3.6e-8)."""
    pr = make_problem(kernel, 3000, p, ncrit=32, dist="cube", seed=1)
    _fmm_vs_direct(pr, p, tol)


@pytest.mark.parametrize("kernel,p,dist,depth,tol", [("laplace", 6, "cube", 3, 2e-5), ("laplace", 6, "clusters", 4, 2e-5),
                                                     ("modified_helmholtz", 5, "sphere", 3, 1e-4)])
def test_oracle_fmm_non_adaptive_against_direct(kernel, p, dist, depth, tol):
    """The same evaluation on the non-adaptive tree (reference NON_ADAPTIVE
    build: uniform to `depth`, ghost empty nodes, target-only child lists)."""
    pr = make_problem(kernel, 3000, p, dist=dist, seed=6, depth=depth)
    _fmm_vs_direct(pr, p, tol)


def _fmm_vs_direct(pr, p, tol):
    orc = O.Oracle(pr["kid"], p, pr["flat"], pr["ops"], pr["wavek"])
    orc.set_charges(pr["qs"])
    v = orc.evaluate()
    tc = pr["flat"]["trg_coord"].reshape(-1, 3)
    idx = np.arange(0, tc.shape[0], 11)
    d = O.direct(pr["kid"], pr["flat"]["src_coord"].reshape(-1, 3), pr["qs"], tc[idx], pr["wavek"])
    assert rel_err(v[idx, 0], d[:, 0]) < tol
    assert rel_err(v[idx, 1:], d[:, 1:]) < 5 * tol


@pytest.mark.parametrize("n", [2, 3, 5, 8, 10, 12, 16, 20, 24, 30])
@pytest.mark.parametrize("sign", [-1, 1])
def test_mixed_radix_fft_matches_numpy(n, sign):
    """The CPU baseline's own FFT (oracle/fast.c) against numpy.fft (radix 2, 3, 4, 5 and their mixes)."""
    x = np.random.default_rng(n).normal(size=n) + 1j * np.random.default_rng(n + 1).normal(size=n)
    ref = np.fft.fft(x) if sign < 0 else np.fft.ifft(x) * n
    assert rel_err(O.fft(x, sign), ref) < 1e-14


@pytest.mark.parametrize("p,dist,n", [(4, "cube", 4000), (5, "plummer", 3000), (6, "clusters", 5000),
                                      (8, "cube", 6000)])
def test_fast_cpu_evaluation_equals_oracle(p, dist, n):
    """The FFT-based CPU evaluation (bench.py's CPU baseline) against the plain
    restatement on the same tree and operators, adaptive trees with P2L / M2P
    lists included (odd p exercises the GEMM's column tail)."""
    pr = make_problem("laplace", n, p, ncrit=32, dist=dist, seed=4, charge="u11")
    orc = O.Oracle(pr["kid"], p, pr["flat"], pr["ops"])
    orc.set_charges(pr["qs"])
    ref = orc.evaluate().copy()
    up, dn = orc.up.copy(), orc.dn.copy()
    v = orc.evaluate_fast()
    assert rel_err(v, ref) < 1e-12
    # equivalent densities come out of the ill-conditioned pseudo-inverses,
    # which amplify last-bit differences of the check potentials (other
    # summation order in the FFT and the GEMMs) into near-null directions:
    # compared by the check potentials they generate
    f = pr["flat"]
    lev = np.repeat(np.arange(len(f["level_offset"]) - 1), np.diff(f["level_offset"]))
    rows = np.nonzero((f["src_count"] > 0) | (f["child0"] >= 0))[0]
    assert equiv_rel_diff(0, 0.0, p, f["r0"], lev, orc.up, up, rows, True) < 1e-12
    leaf = np.nonzero((f["child0"] < 0) & (f["trg_count"] > 0))[0]
    assert equiv_rel_diff(0, 0.0, p, f["r0"], lev, orc.dn, dn, leaf, False) < 1e-12
