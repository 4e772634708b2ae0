"""Benchmark: Laplace FMM evaluation, N=2M uniform, p=10, ncrit=128 on one MI355X.

One step = upward_pass + downward_pass of the whole tree with charges already
resident in HBM (the FMM evaluation phase, every operator including the
near-field P2P, which each upward_pass starts anew on a second stream; the
host tree/list/precompute setup is reported separately).  The headline `value` is the mixed-precision
path of this workload: FP32 near-field P2P (Kahan-combined 64-source tiles), FP32 M2L (transforms, and the
Hadamard products on the f16 matrix cores with split-FP16 operands and FP32 accumulation), FP32 check
potentials, L2P / M2P evaluation and M2M / L2L translations, with the check-to-equivalent pseudo-inverse
solves of P2M and L2P, the operators and all storage in FP64; its pseudo-inverses drop the singular values
below 4e-8 max(S).  Timed beside it: the FLOAT mode (the counterpart of the reference's -DFLOAT build: the
solves in FP32 too) and the plain FP64 path (untruncated operators).  Errors are relative L2 against an FP64
direct sum on 2,000 seeded sample targets computed by the engine's direct-sum kernel.  The CPU baseline is
oracle/fast.c (the FFT-based CPU evaluation) on the same tree and operators, timed on the host cores.

    python bench.py [--gpus 1] [--steps 5] [--warmup 2] [--n 2000000] [--no-cpu-baseline]

The other workloads of the same evaluation are measured with --workload (one JSON line each, not the
headline): laplace_small (N=100k, p=8, ncrit=64, charges U[0,1)), yukawa_sphere (k=5, N=1M on the
unit sphere, charges N(0,1), p=8, ncrit=64), helmholtz_cube (k=20, N=4M uniform, complex charges,
p=10, ncrit=128) and laplace_batched (64 Plummer problems, N log-uniform in [1e4, 5e5], charges 1/N,
p=8, ncrit=64; all problems merged into one forest, one launch per operator, the launch shapes picked by
the autotuner, one captured graph replayed per step).
"""
import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "exafmm-t_amd"))

METRIC = "particles/s, Laplace FMM N=2M at err≤1e-5: mixed-precision vs FP64 on GPU"

# algorithmic counts (fixed before measurement)
P2P_FLOP_PER_PAIR = 20          # Laplace potential + gradient, rsqrt = 1
HAD_FLOP_PER_PAIR_FREQ = 512    # 8x8 complex matrix x 8-vector per (src, trg) pair per frequency
PEAK = {"fp32_vector": 157.3e12, "fp64_vector": 78.6e12, "fp32_mfma": 157.3e12, "fp64_mfma": 78.6e12,
        "fp16_mfma": 2.5e15, "hbm": 8.0e12}


DTYPE = {"mixed": "f32 P2P (Kahan across tiles), f32 M2L (transforms; Hadamard on f16 matrix cores with split-f16 "
                  "operands and f32 accumulation), f32 check potentials, L2P evaluation and M2M / L2L; f64 "
                  "pseudo-inverse solves of P2M / L2P (pseudo-inverses truncated at 4e-8 max(S)); f64 operators "
                  "and storage",
         "float": "f32 arithmetic in every evaluation stage incl. the pseudo-inverse solves (f16-split Hadamard, "
                  "f32 MFMA GEMMs), f64 storage and setup; pseudo-inverses truncated at 4e-8 max(S)",
         "fp64": "f64"}
PRECISION_LABEL = {"mixed": "mixed: f32 P2P, M2L, check potentials, L2P, M2M/L2L + f64 pseudo-inverse solves",
                   "float": "float: every evaluation stage f32", "fp64": "fp64"}
PATHS = (("mixed", 1), ("float", 2), ("fp64", 0))  # abi.MIXED, abi.FLOAT, abi.FP64; the first is the headline


def compare_fields(res, head):
    """Extra fields of a JSON line: every path other than the headline (value, ms, errors) and the
    headline / FP64 throughput ratio."""
    out = {}
    for name, r in res.items():
        if name == head:
            continue
        out[name + "_value"] = r["value"]
        out[name + "_ms_per_step"] = r["ms"]
        out[name + "_err_phi"], out[name + "_err_grad"] = r["err"][0], r["err"][1]
        if "kernels" in r:
            out[name + "_kernels_ms"] = r["kernels"]
    if "fp64" in res and head != "fp64":
        out["mixed_vs_fp64"] = res[head]["value"] / res["fp64"]["value"]
    return out


def had16_active():
    """True when the mixed evaluation runs its Hadamard on split FP16 (the default stage set)."""
    st = os.environ.get("XFMM_MIXED_STAGES", "p2p,m2l,had16,fft")
    return "m2l" in st and "had16" in st


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def setup_problem(n, p, ncrit, seed):
    import exafmm
    L = exafmm.laplace
    rng = np.random.default_rng(seed)
    x = rng.random((n, 3))
    q = rng.uniform(-1.0, 1.0, n)
    fmm = L.LaplaceFmm(p, ncrit, os.path.join(tempfile.gettempdir(), "bench_laplace_p%d.dat" % p))
    t0 = time.perf_counter()
    tree = L.setup(L.init_sources(x, q), L.init_targets(x), fmm, precompute=False)
    t_setup = time.perf_counter() - t0
    t0 = time.perf_counter()
    flat = fmm.flat_tree(tree)
    qs = fmm.sorted_charges(tree)
    t_flat = time.perf_counter() - t0
    t0 = time.perf_counter()
    fmm.precompute()
    ops = fmm.operators()
    t_pre = time.perf_counter() - t0
    return dict(x=x, q=q, flat=flat, ops=ops, ops_trunc=truncated_operators(fmm), qs=qs, depth=fmm.depth,
                t_setup=t_setup, t_flat=t_flat, t_pre=t_pre)


def truncated_operators(fmm):
    """Operators of the mixed and FLOAT modes: the same precompute with the pseudo-inverses truncated at
    4e-8 max(S)."""
    from exafmm import abi
    save = fmm.precision
    fmm.precision = abi.MIXED
    fmm.precompute()
    ops = fmm.operators()
    fmm.precision = save
    fmm.precompute()
    return ops


def operators_for(prob, precision):
    from exafmm import abi
    return prob["ops"] if precision == abi.FP64 else prob["ops_trunc"]


def p2p_pairs(flat):
    off, src = flat["p2p_offset"], flat["p2p_source"]
    cnt = flat["src_count"]
    per_target = np.add.reduceat(cnt[src], off[:-1]) if src.size else np.zeros(flat["nnodes"])
    per_target = np.where(off[1:] > off[:-1], per_target, 0)
    return int(np.sum(per_target.astype(np.int64) * flat["trg_count"]))


def run_path(prob, p, precision, steps, warmup, graph=True):
    """Times `steps` evaluations (upward + downward pass) after `warmup` eager ones.  With graph=True a
    step replays the evaluation captured as one HIP graph (the same kernels, without the per-operator
    timing events and launch gaps); the per-operator times come from one extra eager, event-timed
    evaluation before the timed region."""
    from exafmm import abi
    eng = abi.Engine(abi.LAPLACE, p, 0.0, precision, 0)
    eng.set_operators(operators_for(prob, precision))
    eng.set_tree(prob["flat"])
    eng.set_charges(prob["qs"])
    eng.synchronize()
    for _ in range(warmup):
        eng.upward_pass()
        eng.downward_pass()
    eng.synchronize()
    eng.upward_pass()
    eng.downward_pass()
    eng.synchronize()
    kt = eng.timings()
    if graph:
        eng.graph_capture()
        eng.graph_launch()
        eng.synchronize()

    def step():
        if graph:
            eng.graph_launch()
        else:
            eng.upward_pass()
            eng.downward_pass()

    per_step = []
    t0 = time.perf_counter()
    for _ in range(steps):
        ts = time.perf_counter()
        step()
        eng.synchronize()
        per_step.append(time.perf_counter() - ts)
    total = time.perf_counter() - t0
    vals = eng.values()
    eng.close()
    return dict(total=total, per_step=per_step, kernels=kt, values=vals, graph=graph)


def sample_error(prob, values, nsample, seed):
    from exafmm import abi
    flat = prob["flat"]
    tc = flat["trg_coord"].reshape(-1, 3)
    rng = np.random.default_rng(seed)
    idx = np.sort(rng.choice(tc.shape[0], size=min(nsample, tc.shape[0]), replace=False))
    ref = abi.direct(abi.LAPLACE, flat["src_coord"].reshape(-1, 3), prob["qs"], tc[idx], 0.0, abi.FP64, True)
    v = values[idx]
    ep = float(np.linalg.norm(v[:, 0] - ref[:, 0]) / np.linalg.norm(ref[:, 0]))
    eg = float(np.linalg.norm(v[:, 1:] - ref[:, 1:]) / np.linalg.norm(ref[:, 1:]))
    return ep, eg


def omp_threads():
    """Threads the OpenMP oracle runs with: OMP_NUM_THREADS when set, else the CPUs this process may run on
    (its affinity mask, not the machine's core count)."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    avail = len(os.sched_getaffinity(0))
    return min(int(env), avail) if env.isdigit() and int(env) > 0 else avail


def cpu_baseline(prob, p, ncrit, n):
    """The CPU baseline: the oracle's FFT-based evaluation (oracle/fast.c: the reference's algorithm at its
    cost -- FFT transforms, blocked Hadamard, GEMM translations, vectorised P2P; FP64, OpenMP on the host
    cores) on the same tree, charges and FP64 operators as the GPU paths, at the full N: one untimed and
    one timed evaluation."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    orc = O.Oracle(O.LAPLACE, p, prob["flat"], prob["ops"])
    orc.set_charges(prob["qs"])
    orc.evaluate_fast()
    t0 = time.perf_counter()
    orc.evaluate_fast()
    dt = time.perf_counter() - t0
    st = {k: round(v, 3) for k, v in orc.fast_stage_s.items()}
    cores = omp_threads()
    return {"value": n / dt, "unit": "particles/s", "cores": cores, "kind": "port",
            "sample": "oracle/fast.c full upward+downward pass of the same workload (N=%d uniform, p=%d, "
                      "ncrit=%d, FP64), %.2f s; stages (s) %s" % (n, p, ncrit, dt, json.dumps(st))}


def kernel_source_sha256(name, path=None):
    """sha256 of the text of the kernel `name` (from its `void name(` line to the closing brace at column
    0) in the engine's source: the key that ties a committed counter measurement to the code it measured."""
    import hashlib
    path = path or os.path.join(ROOT, "exafmm-t_amd", "csrc", "m2l.hip")
    with open(path) as f:
        src = f.read()
    i = src.find("void %s(" % name)
    if i < 0:
        return None
    j = src.find("\n}\n", i)
    return hashlib.sha256(src[i:j + 3].encode()).hexdigest()


def measured_traffic(dom, mixed, a):
    """HBM bytes per launch of the dominant kernel from the committed FETCH_SIZE / WRITE_SIZE
    measurement of the same workload (profiles/), scaled by the copy-kernel calibration; None when
    no measurement matches this kernel, its current source text and the workload."""
    path = os.path.join(ROOT, "profiles", "laplace_N%s_p%d_hadamard_traffic.json" % (
        "2M" if a.n == 2_000_000 else str(a.n), a.p))
    if dom != "HAD" or a.ncrit != 128 or not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    want = ("k_hadamard_h" if had16_active() else "k_hadamard<float>") if mixed else "k_hadamard<double>"
    if t.get("kernel") != want:
        return None
    if t.get("kernel_source_sha256") != kernel_source_sha256(want.split("<")[0]):
        return None
    c = t["calibration"]
    return 1024.0 * (t["fetch_kib_per_dispatch"] * c["fetch_scale"] + t["write_kib_per_dispatch"] * c["write_scale"])


# ------------------------------------------------------------------ other workloads

def make_points(dist, n, rng):
    if dist == "cube":
        return rng.random((n, 3))
    if dist == "sphere":  # normalised Gaussian samples on the unit sphere
        x = rng.normal(size=(n, 3))
        return x / np.linalg.norm(x, axis=1)[:, None]
    if dist == "plummer":  # scale radius 1, radius truncated at 10
        m = rng.random(n) * 0.999
        r = np.minimum(1.0 / np.sqrt(m ** (-2.0 / 3.0) - 1.0), 10.0)
        v = rng.normal(size=(n, 3))
        return v / np.linalg.norm(v, axis=1)[:, None] * r[:, None]
    raise ValueError(dist)


WORKLOADS = {
    # name: kernel id, N, p, ncrit, wavek, distribution, charges
    "laplace_small": (0, 100_000, 8, 64, 0.0, "cube", "u01"),
    "yukawa_sphere": (2, 1_000_000, 8, 64, 5.0, "sphere", "normal"),
    "helmholtz_cube": (1, 4_000_000, 10, 128, 20.0, "cube", "complex"),
}


def setup_generic(kid, n, p, ncrit, wavek, dist, charges, seed):
    import exafmm
    mod = [exafmm.laplace, exafmm.helmholtz, exafmm.modified_helmholtz][kid]
    rng = np.random.default_rng(seed)
    x = make_points(dist, n, rng)
    if charges == "u01":
        q = rng.random(n)
    elif charges == "normal":
        q = rng.normal(size=n)
    elif charges == "complex":
        q = rng.uniform(-1.0, 1.0, n) + 1j * rng.uniform(-1.0, 1.0, n)
    else:
        q = np.full(n, float(charges))
    fname = os.path.join(tempfile.gettempdir(), "bench_k%d_p%d_w%g.dat" % (kid, p, wavek))
    if kid == 0:
        fmm = mod.LaplaceFmm(p, ncrit, fname)
    elif kid == 1:
        fmm = mod.HelmholtzFmm(p, ncrit, complex(wavek), fname)
    else:
        fmm = mod.ModifiedHelmholtzFmm(p, ncrit, wavek, fname)
    t0 = time.perf_counter()
    tree = mod.setup(mod.init_sources(x, q), mod.init_targets(x), fmm)
    flat = fmm.flat_tree(tree)
    ops = fmm.operators()
    qs = fmm.sorted_charges(tree)
    return dict(kid=kid, p=p, wavek=wavek, flat=flat, ops=ops, ops_trunc=truncated_operators(fmm), qs=qs,
                depth=fmm.depth, n=n, t_setup=time.perf_counter() - t0)


def engine_for(prob, precision):
    from exafmm import abi
    eng = abi.Engine(prob["kid"], prob["p"], prob["wavek"], precision, 0)
    eng.set_operators(operators_for(prob, precision))
    eng.set_tree(prob["flat"])
    eng.set_charges(prob["qs"])
    return eng


def sample_error_generic(prob, values, nsample, seed):
    from exafmm import abi
    flat = prob["flat"]
    tc = flat["trg_coord"].reshape(-1, 3)
    rng = np.random.default_rng(seed)
    idx = np.sort(rng.choice(tc.shape[0], size=min(nsample, tc.shape[0]), replace=False))
    ref = abi.direct(prob["kid"], flat["src_coord"].reshape(-1, 3), prob["qs"], tc[idx], prob["wavek"], abi.FP64,
                     True)
    v = values[idx]
    ep = float(np.linalg.norm(v[:, 0] - ref[:, 0]) / np.linalg.norm(ref[:, 0]))
    eg = float(np.linalg.norm(v[:, 1:] - ref[:, 1:]) / np.linalg.norm(ref[:, 1:]))
    return ep, eg


def run_single_workload(a, name):
    from exafmm import abi
    kid, n, p, ncrit, wavek, dist, charges = WORKLOADS[name]
    log("setup %s: N=%d p=%d ncrit=%d" % (name, n, p, ncrit))
    prob = setup_generic(kid, n, p, ncrit, wavek, dist, charges, seed=0)
    log("setup %.1fs depth %d nodes %d" % (prob["t_setup"], prob["depth"], prob["flat"]["nnodes"]))
    res = {}
    for pname, prec in PATHS:
        eng = engine_for(prob, prec)
        for _ in range(a.warmup):
            eng.upward_pass()
            eng.downward_pass()
        eng.synchronize()
        t0 = time.perf_counter()
        for _ in range(a.steps):
            eng.upward_pass()
            eng.downward_pass()
        eng.synchronize()
        dt = (time.perf_counter() - t0) / a.steps
        kt = eng.timings()
        err = sample_error_generic(prob, eng.values(), 1000, seed=123)
        eng.close()
        res[pname] = dict(ms=1e3 * dt, value=n / dt, err=err, kernels={k: round(v, 4) for k, v in kt.items()})
        log("%s %s: %.3f ms/step err phi %.2e grad %.2e" % (name, pname, 1e3 * dt, err[0], err[1]))
    m = res["mixed"]
    line = {
        "metric": "particles/s, %s FMM evaluation" % name, "value": m["value"], "unit": "particles/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": m["ms"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": DTYPE["mixed"] + (" (complex)" if kid == 1 else ""),
        "data": "synthetic (seeded %s, charges %s)" % (dist, charges),
        "config": {"workload": "%s_N%d_p%d_ncrit%d" % (name, n, p, ncrit), "kernel": ["laplace", "helmholtz",
                   "modified_helmholtz"][kid], "wavek": wavek, "N": n, "p": p, "ncrit": ncrit,
                   "depth": prob["depth"], "nodes": int(prob["flat"]["nnodes"])},
        "err_phi": m["err"][0], "err_grad": m["err"][1], "kernels_ms": m["kernels"],
        "setup_s": round(prob["t_setup"], 2), "roofline": None, "cpu_baseline": None,
    }
    line["config"]["precision"] = PRECISION_LABEL["mixed"]
    line.update(compare_fields(res, "mixed"))
    return line


def plummer_batch(nprob, seed=5):
    """The batched workload: nprob Plummer problems (scale radius 1, radius truncated at 10), N log-uniform in
    [10,000, 500,000], charges 1/N, targets = sources."""
    rng = np.random.default_rng(seed)
    ns = np.exp(rng.uniform(np.log(1e4), np.log(5e5), nprob)).astype(int)
    probs = []
    for i, n in enumerate(ns):
        x = make_points("plummer", int(n), np.random.default_rng(100 + i))
        probs.append((x, np.full(int(n), 1.0 / float(n)), x))
    return ns, probs


def run_batched(a, nprob=64):
    """nprob independent Laplace problems evaluated as one forest (exafmm.batch): one operator set, one launch
    per operator for all problems, the launch shapes picked by the engine's autotuner, one captured graph
    replayed per step.  Per-problem errors against the FP64 direct sum on 200 sampled targets each."""
    from exafmm import abi
    from exafmm.batch import LaplaceBatch, LaplaceForest
    ns, probs = plummer_batch(nprob)
    t0 = time.perf_counter()
    forest = LaplaceForest(probs, p=8, ncrit=64)
    t_setup = time.perf_counter() - t0
    log("batched setup %.1fs, %d problems, %d particles, forest %d nodes depth %d"
        % (t_setup, nprob, int(ns.sum()), forest.flat["nnodes"], forest.depth))
    fname = os.path.join(tempfile.gettempdir(), "bench_batch_p8.dat")
    res = {}
    for pname, prec in PATHS:
        b = LaplaceBatch(forest, precision=prec, filename=fname, autotune=True, graph=True)
        for _ in range(max(1, a.warmup)):
            b.run()
        b.engine.synchronize()
        t0 = time.perf_counter()
        for _ in range(a.steps):
            b.run()
        b.engine.synchronize()
        dt = (time.perf_counter() - t0) / a.steps
        vals = b.values()
        errs = []
        for i, (x, q, _) in enumerate(probs):
            idx = np.sort(np.random.default_rng(7 + i).choice(x.shape[0], 200, replace=False))
            ref = abi.direct(abi.LAPLACE, x, q, x[idx], 0.0, abi.FP64, True)
            v = vals[i][idx]
            errs.append((float(np.linalg.norm(v[:, 0] - ref[:, 0]) / np.linalg.norm(ref[:, 0])),
                         float(np.linalg.norm(v[:, 1:] - ref[:, 1:]) / np.linalg.norm(ref[:, 1:]))))
        tuning = b.tuning
        b.close()
        res[pname] = dict(ms=1e3 * dt, value=float(ns.sum()) / dt, tuning=tuning,
                          err=(max(x[0] for x in errs), max(x[1] for x in errs)))
        log("batched %s: %.3f ms per batch, max err phi %.2e grad %.2e, tuning %s"
            % (pname, 1e3 * dt, res[pname]["err"][0], res[pname]["err"][1], tuning))
    m = res["mixed"]
    line = {
        "metric": "particles/s, laplace_batched FMM evaluation (aggregate)", "value": m["value"],
        "unit": "particles/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": m["ms"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": DTYPE["mixed"],
        "data": "synthetic (seeded Plummer, scale radius 1, truncated at r=10; charges 1/N)",
        "config": {"workload": "laplace_batched_%dproblems_p8_ncrit64" % nprob, "kernel": "laplace",
                   "problems": nprob, "N_total": int(ns.sum()), "N_min": int(ns.min()), "N_max": int(ns.max()),
                   "p": 8, "ncrit": 64, "forest_nodes": int(forest.flat["nnodes"]), "depth": forest.depth,
                   "graph": "one captured evaluation of the whole forest", "tuning": m["tuning"],
                   "precision": PRECISION_LABEL["mixed"]},
        "err_phi": m["err"][0], "err_grad": m["err"][1],
        "setup_s": round(t_setup, 2), "roofline": None, "cpu_baseline": None,
    }
    line.update(compare_fields(res, "mixed"))
    return line


def roofline(dom, ms, mixed, h16, npairs, m2l_pairs, nfreq, nsrc, ntrg):
    """Roofline of the dominant kernel `dom` ("P2P", "HAD", "FFT" or "IFFT") of
    one evaluation that took `ms` milliseconds.  For the Hadamard the bound is
    min(MFMA peak, AI x HBM bandwidth): whichever of the two times
    (algorithmic flops / peak, algorithmic bytes / HBM peak) is longer names
    it, and the other view is reported beside it.  `traffic` is left to the
    caller (measured counters)."""
    t = ms * 1e-3
    rb = 4 if mixed else 8
    if dom == "P2P":
        flops = P2P_FLOP_PER_PAIR * npairs
        roof = {"kernel": "P2P near field (k_p2p_smem, %s)" % ("FP32" if mixed else "FP64"),
                "bound": "valu", "achieved": flops / t / 1e12,
                "peak": (PEAK["fp32_vector"] if mixed else PEAK["fp64_vector"]) / 1e12, "unit": "TFLOP/s",
                "traffic": None, "algorithmic": "%d pairs x %d flop" % (npairs, P2P_FLOP_PER_PAIR)}
    elif dom == "HAD":
        flops = HAD_FLOP_PER_PAIR_FREQ * m2l_pairs * nfreq
        name = "k_hadamard_h, split FP16 on the f16 MFMA (3 f16 products per FP32-equivalent product)" \
            if h16 else "k_hadamard, %s" % ("FP32" if mixed else "FP64")
        peak = PEAK["fp16_mfma"] / 3 if h16 else (PEAK["fp32_mfma"] if mixed else PEAK["fp64_mfma"])
        # Laplace is scale-invariant: one set of 26 operators serves every level
        byts = (nsrc + ntrg) * nfreq * 8 * 2 * rb + 26 * nfreq * 128 * rb
        mfma_view = {"achieved": flops / t / 1e12, "peak": peak / 1e12, "unit": "TFLOP/s",
                     "frac": flops / t / peak,
                     "peak_note": "f16 MFMA 2500 TFLOP/s / 3 products" if h16 else "dense MFMA peak"}
        hbm_view = {"achieved": byts / t / 1e9, "peak": PEAK["hbm"] / 1e9, "unit": "GB/s",
                    "frac": byts / t / PEAK["hbm"]}
        alg = "%d pairs x %d freq x %d flop; %d source + %d target nodes x %d freq x 8 children complex " \
              "(%d B reals) + 26 operators x %d freq x 128 reals" % (
                  m2l_pairs, nfreq, HAD_FLOP_PER_PAIR_FREQ, nsrc, ntrg, nfreq, rb, nfreq)
        if byts / PEAK["hbm"] >= flops / peak:
            roof = dict(hbm_view, bound="hbm", kernel="M2L hadamard_product (%s)" % name, traffic=None,
                        algorithmic=alg, algorithmic_bytes=byts, mfma=mfma_view)
        else:
            roof = dict(mfma_view, bound="mfma", kernel="M2L hadamard_product (%s)" % name, traffic=None,
                        algorithmic=alg, algorithmic_flops=flops, hbm=hbm_view)
    else:
        nslots = nsrc if dom == "FFT" else ntrg
        byts = nslots * nfreq * 8 * 2 * rb
        roof = {"kernel": "M2L %s" % dom, "bound": "hbm", "achieved": byts / t / 1e9,
                "peak": PEAK["hbm"] / 1e9, "unit": "GB/s", "traffic": None,
                "algorithmic": "%d nodes x %d freq x 8 children complex" % (nslots, nfreq)}
    roof["frac"] = roof["achieved"] / roof["peak"]
    roof["ms"] = ms
    return roof


def pinv_rank(ops, key, nsurf):
    """Rank of a truncated pseudo-inverse factor: 1 + the last column of C2E_V (nsurf x nsurf per level,
    row-major) that is not exactly zero, the number of singular directions the engine applies."""
    v = np.asarray(ops[key]).reshape(-1, nsurf, nsurf)
    nz = np.nonzero(np.any(v != 0.0, axis=(0, 1)))[0]
    return int(nz[-1]) + 1 if nz.size else 1


def evaluation_roofline(flat, ops, p, precision, kernels_ms, ms, npairs, m2l_pairs, nfreq, nsrc, ntrg):
    """Roofline of the whole evaluation: for every stage its own roofline time (algorithmic flops / the
    peak of the unit it runs on, or algorithmic bytes / HBM peak, whichever is longer), their sum (a
    serial lower bound: every stage fills the chip on its own), and that sum over the measured time per
    step.  Precision per stage follows the path: FP64 everywhere, or the mixed policy (FP32 near field,
    check potentials, L2P evaluation, M2M / L2L, transforms; split-FP16 Hadamard; FP64 pseudo-inverse
    solves), or FLOAT (every stage FP32)."""
    nsurf = 6 * (p - 1) ** 2 + 2
    leaf = flat["child0"] < 0
    nleaf_s = int(np.sum(leaf & (flat["src_count"] > 0)))
    nleaf_t = int(np.sum(leaf & (flat["trg_count"] > 0)))
    nonleaf = int(np.sum(~leaf))
    ns, nt = int(np.sum(flat["src_count"])), int(np.sum(flat["trg_count"]))
    r_up, r_dn = pinv_rank(ops, "uc2e_v", nsurf), pinv_rank(ops, "dc2e_v", nsurf)
    f32 = precision != 0
    valu = PEAK["fp32_vector"] if f32 else PEAK["fp64_vector"]
    trans = PEAK["fp32_mfma"] if f32 else PEAK["fp64_mfma"]
    pinv = PEAK["fp32_mfma"] if precision == 2 else PEAK["fp64_mfma"]
    rb = 4 if f32 else 8
    had_peak = PEAK["fp16_mfma"] / 3 if f32 else PEAK["fp64_mfma"]
    # stage -> kernels -> views (unit, seconds); a kernel takes its longest view, a stage the sum of its kernels
    stages = {
        # check potentials (12 flop per pair), then the two pseudo-inverse GEMMs over r directions
        "P2M": [[("valu", 12.0 * ns * nsurf / valu)], [("mfma", 4.0 * nleaf_s * nsurf * r_up / pinv)]],
        "M2M": [[("mfma", 16.0 * nonleaf * nsurf * nsurf / trans)]],
        "P2P": [[("valu", float(P2P_FLOP_PER_PAIR) * npairs / valu)]],
        "FFT": [[("hbm", (nsrc * nfreq * 8 * 2 * rb + nsrc * 8 * nsurf * 8) / PEAK["hbm"])]],
        "HAD": [[("hbm", ((nsrc + ntrg) * nfreq * 8 * 2 * rb + 26 * nfreq * 128 * rb) / PEAK["hbm"]),
                 ("mfma", float(HAD_FLOP_PER_PAIR_FREQ) * m2l_pairs * nfreq / had_peak)]],
        "IFFT": [[("hbm", (ntrg * nfreq * 8 * 2 * rb + 2 * ntrg * 8 * nsurf * 8) / PEAK["hbm"])]],
        "L2L": [[("mfma", 16.0 * nonleaf * nsurf * nsurf / trans)]],
        "L2P": [[("mfma", 4.0 * nleaf_t * nsurf * r_dn / pinv)],
                [("valu", float(P2P_FLOP_PER_PAIR) * nt * nsurf / valu)]],
    }
    out, total = {}, 0.0
    for name, kernels in stages.items():
        best = [max(views, key=lambda q: q[1]) for views in kernels]
        t = sum(x for _, x in best)
        out[name] = {"bound": "+".join(b for b, _ in best), "roofline_ms": round(1e3 * t, 4),
                     "ms": kernels_ms.get(name)}
        total += t
    return {"serial_roofline_ms": round(1e3 * total, 4), "ms": ms, "frac": 1e3 * total / ms,
            "pinv_rank": [r_up, r_dn], "stages": out,
            "note": "per stage: algorithmic flops / peak of its unit or bytes / 8 TB/s; FP32 vector peak 157.3, "
                    "FP64 78.6, FP32 MFMA 157.3, FP64 MFMA 78.6, split-FP16 Hadamard 2500/3 TFLOP/s; the FFT "
                    "stage's ms covers the near field running beside it on the second stream"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=2_000_000)
    ap.add_argument("--p", type=int, default=10)
    ap.add_argument("--ncrit", type=int, default=128)
    ap.add_argument("--nsample", type=int, default=2000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--fp64-only", action="store_true")
    ap.add_argument("--mixed-only", action="store_true", help="time the headline mixed-precision path only")
    ap.add_argument("--eager", action="store_true",
                    help="time eager launches (with per-operator events) instead of replays of the captured graph")
    ap.add_argument("--workload", default="laplace_uniform",
                    choices=["laplace_uniform", "laplace_batched"] + sorted(WORKLOADS))
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build evaluates one FMM on one GPU; multi-GPU runs are not part of it "
              "(requested --gpus %d)" % a.gpus, file=sys.stderr)
        sys.exit(2)
    from exafmm import abi
    if abi.device_count() < 1:
        print("bench.py: no GPU visible to the engine", file=sys.stderr)
        sys.exit(1)

    if a.workload == "laplace_batched":
        print(json.dumps(run_batched(a)), flush=True)
        return
    if a.workload != "laplace_uniform":
        print(json.dumps(run_single_workload(a, a.workload)), flush=True)
        return

    log("setup: N=%d p=%d ncrit=%d" % (a.n, a.p, a.ncrit))
    prob = setup_problem(a.n, a.p, a.ncrit, seed=0)
    flat = prob["flat"]
    log("setup %.2fs flatten %.2fs precompute %.2fs depth %d nodes %d"
        % (prob["t_setup"], prob["t_flat"], prob["t_pre"], prob["depth"], flat["nnodes"]))
    npairs = p2p_pairs(flat)
    m2l_pairs = int(np.sum(flat["m2l_source"] >= 0))
    n1 = 2 * a.p
    nfreq = n1 * n1 * (n1 // 2 + 1)

    res = {}
    paths = [("fp64", abi.FP64)] if a.fp64_only else list(PATHS)
    if a.mixed_only:
        paths = [PATHS[0]]
    for name, prec in paths:
        log("timing %s path: warmup %d steps %d" % (name, a.warmup, a.steps))
        r = run_path(prob, a.p, prec, a.steps, a.warmup, graph=not a.eager)
        r["err"] = sample_error(prob, r["values"], a.nsample, seed=123)
        log("%s: %.3f ms/step err phi %.2e grad %.2e kernels %s" % (
            name, 1e3 * r["total"] / a.steps, r["err"][0], r["err"][1],
            {k: round(v, 3) for k, v in r["kernels"].items()}))
        del r["values"]
        r["value"] = a.n / (r["total"] / a.steps)
        r["ms"] = 1e3 * r["total"] / a.steps
        r["kernels"] = {k: round(v, 4) for k, v in r["kernels"].items()}
        res[name] = r

    hname = paths[0][0]
    head = res[hname]
    ms_step = 1e3 * head["total"] / a.steps
    value = a.n / (head["total"] / a.steps)
    kt = head["kernels"]
    # dominant kernel and its roofline among the kernels that run alone (the
    # near field P2P on the second stream and the forward transform on the
    # main stream run at the same time, so neither event interval is one
    # kernel's duration)
    cand = {"HAD": kt.get("HAD", 0.0), "IFFT": kt.get("IFFT", 0.0)}
    dom = max(cand, key=cand.get)
    mixed = hname != "fp64"
    nsrc = int(np.unique(flat["m2l_source"][flat["m2l_source"] >= 0]).size)
    ntrg = int(np.sum(np.any(flat["m2l_source"].reshape(-1, 26) >= 0, axis=1)))
    roof = roofline(dom, cand[dom], mixed, mixed and had16_active(), npairs, m2l_pairs, nfreq, nsrc, ntrg)
    roof["traffic"] = measured_traffic(dom, mixed, a)

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "particles/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPE[hname],
        "data": "synthetic (seeded uniform [0,1]^3, charges U[-1,1])",
        "config": {"workload": "laplace_fmm_N%d_p%d_ncrit%d_uniform" % (a.n, a.p, a.ncrit), "kernel": "laplace",
                   "N": a.n, "p": a.p, "ncrit": a.ncrit, "depth": prob["depth"], "nodes": int(flat["nnodes"]),
                   "p2p_pairs": npairs, "m2l_pairs": m2l_pairs,
                   "precision": PRECISION_LABEL[hname]},
        "timing": ("eager launches" if a.eager else "replays of the evaluation captured as one HIP graph") +
                  "; kernels_ms from one event-timed eager evaluation",
        "err_phi": head["err"][0],
        "err_grad": head["err"][1],
        "kernels_ms": kt,
        "setup_s": round(prob["t_setup"] + prob["t_flat"], 3), "precompute_s": round(prob["t_pre"], 3),
        "roofline": roof,
        "evaluation_roofline": evaluation_roofline(flat, operators_for(prob, paths[0][1]), a.p, paths[0][1], kt,
                                                   ms_step, npairs, m2l_pairs, nfreq, nsrc, ntrg),
    }
    line.update(compare_fields(res, hname))
    if not a.no_cpu_baseline:
        log("cpu baseline (oracle/fast.c) on the same workload, N=%d" % a.n)
        line["cpu_baseline"] = cpu_baseline(prob, a.p, a.ncrit, a.n)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
