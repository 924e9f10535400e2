"""Benchmark of the variable-block sparse hot path on one MI355X.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu] [--quick]

Workload (the configuration the metric is quoted on): graphene-like
tight-binding Hamiltonian, periodic honeycomb lattice of 200 x 200 cells, 4
orbitals per atom (4 x 4 blocks, N = 320,000 rows, 800,000 blocks), on-site +
nearest + next-nearest neighbour blocks with N(0,1) entries / distance,
spectrum scaled to [-0.99, 0.99]; KPM density of states with 200 moments,
Jackson kernel, 10 Rademacher vectors as one block of columns.

A step is one full KPM moment computation (200 moments x 10 vectors) with the
matrix and vectors resident in HBM.  ``value`` is the FP32-mixed throughput
(FP32 recursion and matrix values, FP64 moment reductions, fused kernel, two
moments per matvec) in moment-vecs/s; the same line carries the plain FP64
GPU path (separate SpMM / axpby / dot kernels, one moment per matvec), the
fused FP64 path, the DOS error of FP32 against FP64, block-SpMV bandwidth on a
1.5 GB variable-block matrix (beyond the 256 MB Infinity Cache), the batched
many-Hamiltonian KPM, the roofline of the dominant kernel and the reference
CPU path timed on this host.
"""
from __future__ import annotations

import argparse
import json
import subprocess
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, published


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    p.add_argument("--quick", action="store_true", help="smaller lattice (smoke of the bench itself)")
    p.add_argument("--only", default="", help="comma list: kpm,spmv,batched,prop,mcweeny,cpu (default all)")
    p.add_argument("--mcweeny-iters", type=int, default=5,
                   help="McWeeny iterations per precision (the configured 10 fill P to ~0.9 G values and take "
                        "9-40 s per precision; 5 keep the default run short)")
    p.add_argument("--mcweeny-legs", default="fp64,fp64_hermitian,fp32,bf16",
                   help="comma list of McWeeny legs (the first one is the reference of the others' errors)")
    p.add_argument("--batched-plan", default="",
                   help="pin the batched KPM plan as THRESHOLD:PARTS (see kpm.set_batch_plan) instead of autotuning")
    p.add_argument("--batched-legs", default="mixed,fp32,fp64",
                   help="comma list of batched legs; without fp64 the speedups and DOS errors are not computed "
                        "(for counter runs)")
    p.add_argument("--copy-calibration", action="store_true",
                   help="only run a 1 GiB device-to-device copy 5 times (scale for the HBM byte counters)")
    return p.parse_args()


def copy_calibration():
    """A plain copy kernel of known size: 1 GiB read + 1 GiB written per launch.
    rocprofv3's FETCH_SIZE / WRITE_SIZE read on it give the factor between the
    counters and true HBM bytes used for the traffic figure of the roofline."""
    import torch

    n = 1 << 27  # float64 elements = 1 GiB
    a = torch.empty(n, dtype=torch.float64, device="cuda").uniform_()
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read_per_launch": 8 * n, "copy_bytes_written_per_launch": 8 * n, "launches": 5}))


def timed(fn, steps, warmup, sync):
    for _ in range(warmup):
        fn()
    sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    sync()
    return (time.perf_counter() - t0) / steps


def spmm_bytes(A, K, p_val, p_vec, with_z=False, with_w=False):
    """Algorithmic bytes of one SpMM / fused step: values, block indices,
    row pointers, X gathered once, Y written once (+ Z, W reads)."""
    n = A.n
    b = p_val * A.nnz + 4 * A.nnzb + 4 * (A.nb + 1) + p_vec * K * (n + n)
    if with_z:
        b += p_vec * K * n
    if with_w:
        b += p_vec * K * n
    return b


def read_traffic(name):
    """HBM bytes per launch measured by rocprofv3 --pmc (FETCH_SIZE +
    WRITE_SIZE, calibrated on a copy kernel), committed under profiles/."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", open(path).read())
    return float(m.group(1)) if m else None


def host_info():
    """CPU model and topology of the host the CPU baselines run on (lscpu)."""
    info = {}
    try:
        txt = subprocess.run(["lscpu"], stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True,
                             timeout=10).stdout
        for line in txt.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in ("Model name", "CPU(s)", "Socket(s)", "Core(s) per socket", "Thread(s) per core",
                             "NUMA node(s)", "CPU max MHz", "L3"):
                info[k.strip()] = v.strip()
    except Exception:  # lscpu missing: fall back to /proc/cpuinfo
        try:
            models = [ln.split(":", 1)[1].strip() for ln in open("/proc/cpuinfo") if ln.startswith("model name")]
            info = {"Model name": models[0] if models else "?", "CPU(s)": str(len(models))}
        except OSError:
            pass
    info["threads_used"] = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return info


def cpu_baseline(G, K, M, seed):
    """Reference vbcsr C++ core (oracle/_ref) on this host: the ex08 KPM
    composition (mult_dense + scale + axpy + bdot), FP64, OpenMP threads."""
    import oracle

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    if oracle.ref_available():
        m = oracle.Mat.of(G)
        R = oracle.RefMat(m)
        V0 = R.random_mv(K, seed, False)
        R.kpm_moments(V0, 3)  # warm-up: builds the cached apply plan
        mu, secs = R.kpm_moments(V0, M)
        return {"value": M * K / secs, "unit": "moment-vecs/s", "cores": threads, "kind": "reference",
                "sample": f"reference vbcsr C++ core (oracle/_ref, -O3, OpenMP, serial MPI_COMM_NULL) on the "
                          f"same graphene 200x200 Hamiltonian: ex08 KPM composition, {M} moments x {K} vectors, "
                          f"FP64, {secs:.2f} s"}, mu
    m = oracle.Mat.of(G)
    V0 = oracle.rademacher(G.bs, np.arange(G.nb), K, seed, False, G.n)
    Ms = min(M, 12)
    t0 = time.perf_counter()
    mu = oracle.kpm_moments(m, V0, Ms, mode=1)
    secs = time.perf_counter() - t0
    return {"value": Ms * K / secs, "unit": "moment-vecs/s", "cores": 1, "kind": "port",
            "sample": f"CPU restatement (oracle/, 1 thread), {Ms} moments x {K} vectors, FP64, {secs:.2f} s"}, None


def bench_propagation(args, sync, log):
    """C5: differentiable Chebyshev propagation on a 100x100-cell honeycomb
    lattice, 4x4 complex blocks (N = 80,000), Gaussian wave packet of width
    5 cells, 50 steps of dt = 0.5 with 64 Chebyshev terms, loss
    |<psi_target|psi(T)>|^2, gradient with respect to every block value
    (checkpoint every 10 steps), complex128 and complex64; central finite
    differences on 20 random entries of the complex128 gradient."""
    import torch

    from vbcsr_amd import VBCSR
    from vbcsr_amd import autograd as AG
    from vbcsr_amd import workloads as W

    nx = 30 if args.quick else 100
    A = W.graphene(nx, nx, 4, seed=5, complex_blocks=True)
    H = VBCSR.from_block_csr(A.bs, A.row_ptr, A.col_ind, A.vals)
    psi0 = W.wavepacket(nx, nx, 4, 5.0, seed=6)
    rng = np.random.default_rng(7)
    tgt = rng.standard_normal(A.n) + 1j * rng.standard_normal(A.n)
    tgt /= np.linalg.norm(tgt)
    vals = torch.as_tensor(A.vals, device="cuda")
    tg = torch.as_tensor(tgt, device="cuda")
    dt, nsteps, nterms, half = 0.5, 50, 64, 1.0  # spectrum scaled to [-0.99, 0.99]

    def loss_of(out):
        ov = (tg.conj().to(out.dtype) * out).sum()
        return ov.real ** 2 + ov.imag ** 2

    res = {"workload": f"honeycomb {nx}x{nx} cells, 4x4 complex blocks (N={A.n}), {nsteps} steps dt={dt}, "
                       f"{nterms} Chebyshev terms, checkpoint 10"}
    grads = {}
    for prec, cdt in (("fp64", torch.complex128), ("fp32", torch.complex64)):
        p0 = torch.as_tensor(psi0, device="cuda").to(cdt)

        def fwd():
            with torch.no_grad():
                return AG.propagate(H, vals, p0, dt, nsteps, nterms, half_width=half, precision=prec)

        def fwd_bwd():
            v = vals.clone().requires_grad_(True)
            loss_of(AG.propagate(H, v, p0, dt, nsteps, nterms, half_width=half, checkpoint=10,
                                 precision=prec)).backward()
            grads[prec] = v.grad
        tf = timed(fwd, 3, 1, sync)
        tb = timed(fwd_bwd, 2, 1, sync)
        norm = float(torch.linalg.norm(fwd().to(torch.complex128)).item())
        res[prec] = {"forward_ms": tf * 1e3, "forward_backward_ms": tb * 1e3, "norm_psi_T": norm,
                     "spmv_per_s_forward": nsteps * (nterms - 1) / tf}
        log(f"propagation {prec}: forward {tf * 1e3:.1f} ms, forward+backward {tb * 1e3:.1f} ms, "
            f"|psi(T)| = {norm:.12f}")
    g64 = grads["fp64"].cpu().numpy()
    res["grad_rel_err_fp32_vs_fp64"] = float(np.abs(grads["fp32"].cpu().numpy() - g64).max() / np.abs(g64).max())
    p0 = torch.as_tensor(psi0, device="cuda")
    eps, worst = 1e-6, 0.0
    with torch.no_grad():
        for idx in rng.choice(A.nnz, 20, replace=False):
            for unit in (1.0, 1j):
                vp, vm = vals.clone(), vals.clone()
                vp[idx] += eps * unit
                vm[idx] -= eps * unit
                fd = (loss_of(AG.propagate(H, vp, p0, dt, nsteps, nterms, half_width=half)) -
                      loss_of(AG.propagate(H, vm, p0, dt, nsteps, nterms, half_width=half))).item() / (2 * eps)
                an = g64[idx].real if unit == 1.0 else g64[idx].imag
                worst = max(worst, abs(fd - an) / max(abs(an), 1e-3 * np.abs(g64).max()))
    res["grad_fd_max_rel_err_fp64"] = worst
    log(f"propagation gradient: FD max rel err {worst:.2e}, fp32 vs fp64 {res['grad_rel_err_fp32_vs_fp64']:.2e}")
    if not args.no_cpu and not args.quick:
        import oracle

        if oracle.ref_available():
            # the reference's propagation is a Python Chebyshev series over
            # BlockSpMat::mult (vbcsr/utils.py:71-126); its SpMVs alone, timed
            # on the host, bound its forward pass from below
            R = oracle.RefMat(oracle.Mat.of(A))
            t_mult = R.time_mult(np.asarray(psi0, np.complex128).reshape(-1), 50)
            est = t_mult * nsteps * (nterms - 1)
            res["cpu_reference"] = {
                "spmv_ms": t_mult * 1e3, "forward_spmv_seconds_estimate": est,
                "threads": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)), "kind": "reference",
                "sample": "vbcsr C++ core BlockSpMat::mult (oracle/_ref), complex128, mean of 50 calls; times "
                          f"{nsteps} x {nterms - 1} SpMVs of the forward (vector ops not counted)",
                "gpu_fp64_forward_speedup_over_estimate": est / (res["fp64"]["forward_ms"] * 1e-3)}
            log(f"propagation, reference CPU SpMV {t_mult * 1e3:.3f} ms -> forward >= {est:.2f} s")
    return res


def bench_kpm_scaled(args, sync, log, nx=800):
    """The C2 KPM step on a lattice large enough that one step's bytes (about
    1.5 GB at FP32) exceed the 256 MB Infinity Cache several times: graphene
    nx x nx cells, spectrum scaled on the device, 200 moments x 10 vectors;
    FP32 and FP64 fused steps, per-launch time from HIP events."""
    import vbcsr_amd
    from vbcsr_amd import VBCSR
    from vbcsr_amd import kpm as KP
    from vbcsr_amd import workloads as W

    t0 = time.perf_counter()
    G = W.graphene(nx, nx, 4, seed=0, scale=False)
    H = VBCSR.from_block_csr(G.bs, G.row_ptr, G.col_ind, G.vals)
    KP.scale_to_unit(H, 0.99)
    V0 = H.create_multivector(10)
    V0.set_random(42, normalize=False)
    sync()
    log(f"graphene {nx}x{nx} built and scaled in {time.perf_counter() - t0:.1f} s")
    res = {"workload": f"graphene {nx}x{nx} cells, 4 orbitals (N={G.n}), 200 moments x 10 vectors, two moments "
                       f"per matvec"}
    for name, prec, p in (("fp32", "fp32", 4), ("fp64", "fp64", 8)):
        KP.kpm_moments(H, V0, 200, mode="two", precision=prec)
        sync()
        KP.kpm_moments(H, V0, 200, mode="two", precision=prec)
        kms, launches = vbcsr_amd._core.kpm_last_step_time()
        per = kms / max(launches, 1)
        alg = spmm_bytes(G, 10, p, p, with_z=True)
        gbs = alg / (per * 1e-3) / 1e9
        res[name] = {"avg_launch_us": per * 1e3, "alg_bytes_per_launch": alg, "achieved_gbs": gbs,
                     "frac_hbm": gbs / HBM_PEAK_GBS}
        log(f"kpm {nx}x{nx} {name}: {per * 1e3:.1f} us per step, {gbs:.0f} GB/s ({100 * gbs / HBM_PEAK_GBS:.1f}%)")
    res["fp32_over_fp64_step_speed"] = res["fp64"]["avg_launch_us"] / res["fp32"]["avg_launch_us"]
    return res


def mcweeny_reference_cpu(P0, thr, iters, budget, log):
    """The reference C++ core (oracle/_ref, OpenMP) on the same P0: the
    README/ex07 iteration P2 = P.spmm(P), P3 = P2.spmm(P), P = 3 P2 - 2 P3
    (BlockSpMat::spmm and axpby), iterations run until the time budget is
    spent; per-iteration seconds."""
    import oracle

    if not oracle.ref_available():
        return None
    P = oracle.RefMat(oracle.Mat.of(P0))
    per = []
    for it in range(iters):
        t0 = time.perf_counter()
        P2 = P.spmm(P, thr)
        P3 = P2.spmm(P, thr)
        Pn = P2.duplicate()
        Pn.axpby(-2.0, P3, 3.0)  # self = -2 P3 + 3 P2
        P = Pn
        per.append(time.perf_counter() - t0)
        log(f"mcweeny reference CPU iteration {it}: {per[-1]:.2f} s, {P.nnzb} blocks")
        if sum(per) > budget:
            break
    return per


def bench_mcweeny(args, sync, log):
    """C3: McWeeny purification P <- 3P^2 - 2P^3 with thresholded SpGEMM
    (threshold 1e-5), 5000 block rows of sizes 4..8, ~30 neighbours per row,
    P0 = 0.1 H + 0.5 I with H divided by its largest eigenvalue magnitude
    (Lanczos; as written, without it, P0's spectrum leaves [-0.5, 1.5] and the
    iteration diverges), so P0's spectrum spans [0.4, 0.6].  Numeric phase in
    FP64 (full and Hermitian products), FP32 and BF16 (matrix cores);
    idempotency ||P^2 - P||_F and trace in FP64, and each leg's P against the
    FP64 leg's.  The reference C++ core runs the first iterations on the host
    CPU for comparison.  An iteration that passes the time budget ends the run
    for that precision (reported)."""
    from vbcsr_amd import VBCSR
    from vbcsr_amd import workloads as W

    nb = 1000 if args.quick else 5000
    iters, budget, thr = args.mcweeny_iters, 300.0, 1e-5
    P0 = W.mcweeny_start(nb=nb, seed=3, normalize="spectral")
    lo, hi = P0.meta["spectrum"]
    res = {"workload": f"{nb} block rows, blocks 4-8, ~30 neighbours, H / max|eig H| (Lanczos: [{lo:.3f}, {hi:.3f}]),"
                       f" P0 = 0.1 H + 0.5 I, threshold {thr:g}, up to {iters} iterations"}
    P_ref = None
    # "fp64_hermitian": both products by spmm_hermitian (P, P^2 and P^3 are
    # symmetric: only the upper block triangle is multiplied, then mirrored)
    legs = [g for g in args.mcweeny_legs.split(",") if g]
    bad = set(legs) - {"fp64", "fp64_hermitian", "fp32", "bf16"}
    if bad:
        raise SystemExit(f"unknown McWeeny legs {sorted(bad)}")
    for leg in legs:
        numeric = leg.split("_")[0]
        herm = leg.endswith("hermitian")
        P = VBCSR.from_block_csr(P0.bs, P0.row_ptr, P0.col_ind, P0.vals)
        sync()
        t0 = time.perf_counter()
        done = 0
        per_iter, stats = [], []
        for it in range(iters):
            ti = time.perf_counter()
            if herm:
                P2 = P.spmm_hermitian(P, thr, numeric=numeric)
                s1 = P.spmm_stats
                P3 = P2.spmm_hermitian(P, thr, numeric=numeric)
            else:
                P2 = P.spmm(P, thr, numeric=numeric)
                s1 = P.spmm_stats
                P3 = P2.spmm(P, thr, numeric=numeric)
            s2 = P2.spmm_stats
            P2.axpby(-2.0, P3, 3.0)  # P <- 3 P2 - 2 P3 in place of P2 (BlockSpMat::axpby; P2 is not needed again)
            P = P2
            sync()
            per_iter.append(time.perf_counter() - ti)
            stats.append({"symbolic_pairs": int(s1[0] + s2[0]), "numeric_pairs": int(s1[1] + s2[1]),
                          "blocks_P2": int(s1[3]), "blocks_P3": int(s2[3])})
            done += 1
            log(f"mcweeny {leg} iteration {it}: {per_iter[-1] * 1e3:.1f} ms, {P.nnz} values, "
                f"numeric pairs {stats[-1]['numeric_pairs']}")
            if time.perf_counter() - t0 > budget:  # safety stop only; reported as done < iters
                break
        secs = sum(per_iter)  # the iterations only (not the logging between them)
        P2 = P.spmm(P, 0.0)  # FP64 check: ||P^2 - P||_F over the union pattern, and tr P
        idem = float(np.linalg.norm((P2 - P).get_values()))
        del P2
        res[leg] = {"iterations": done, "seconds": secs, "ms_per_iteration": [t * 1e3 for t in per_iter],
                    "spgemm_stats_per_iteration": stats,
                    "trace": float(P.trace()), "values_final": int(P.nnz),
                    "idempotency_frobenius": idem}
        if P_ref is None:
            P_ref = P
        else:
            D = (P - P_ref).get_values()
            res[leg]["vs_fp64"] = {"max_abs": float(np.abs(D).max()) if D.size else 0.0,
                                   "frobenius": float(np.linalg.norm(D)),
                                   "trace_diff": res[leg]["trace"] - res[legs[0]]["trace"], "against": legs[0]}
        log(f"mcweeny {leg}: {done} iterations in {secs:.2f} s, trace {res[leg]['trace']:.6f}, "
            f"||P^2-P||_F {idem:.3e}" + (f", vs fp64 max {res[leg]['vs_fp64']['max_abs']:.2e}"
                                          if "vs_fp64" in res[leg] else ""))
    if not args.no_cpu and not args.quick and "fp64" in res:
        cpu = mcweeny_reference_cpu(P0, thr, iters, 12.0, log)
        if cpu:
            n = len(cpu)
            gpu_same = sum(res["fp64"]["ms_per_iteration"][:n]) / 1e3
            res["cpu_reference"] = {"iterations": n, "seconds": sum(cpu), "s_per_iteration": cpu,
                                    "threads": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)),
                                    "gpu_fp64_seconds_same_iterations": gpu_same,
                                    "gpu_speedup_same_iterations": sum(cpu) / gpu_same if gpu_same > 0 else None}
    return res


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration()
        return
    only = set(filter(None, args.only.split(",")))
    want = lambda k: not only or k in only  # noqa: E731

    import torch

    import vbcsr_amd
    from vbcsr_amd import VBCSR
    from vbcsr_amd import kpm as KP
    from vbcsr_amd import workloads as W

    sync = torch.cuda.synchronize
    nx = 60 if args.quick else 200
    M, K, seed = 200, 10, 42
    out = {}

    log(f"generating graphene {nx}x{nx} x4 orbitals")
    G = W.graphene(nx, nx, 4, seed=0)
    H = VBCSR.from_block_csr(G.bs, G.row_ptr, G.col_ind, G.vals)
    V0 = H.create_multivector(K)
    V0.set_random(seed, normalize=False)
    sync()

    ms_step = None
    roofline = None
    if want("kpm"):
        runs = {}
        for name, mode, prec in (("fp64_plain", "plain", "fp64"), ("fp64_fused", "two", "fp64"),
                                 ("fp32_mixed", "two", "fp32")):
            res = {}

            def step(mode=mode, prec=prec, res=res):
                res["mu"] = KP.kpm_moments(H, V0, M, mode=mode, precision=prec)

            secs = timed(step, args.steps, args.warmup, sync)
            kms, launches = vbcsr_amd._core.kpm_last_step_time()
            runs[name] = dict(secs=secs, mu=res["mu"], kernel_ms=kms, launches=launches)
            log(f"{name}: {secs * 1e3:.3f} ms/step, {M * K / secs:.4g} moment-vecs/s, "
                f"step kernels {kms:.3f} ms / {launches} launches")
        _, rho64 = KP.dos_from_moments(runs["fp64_plain"]["mu"])
        _, rho32 = KP.dos_from_moments(runs["fp32_mixed"]["mu"])
        _, rho64f = KP.dos_from_moments(runs["fp64_fused"]["mu"])
        dos_err = KP.dos_error(rho32, rho64)
        fp32 = runs["fp32_mixed"]
        ms_step = fp32["secs"] * 1e3
        out["value"] = M * K / fp32["secs"]
        out["fp64_plain_moment_vecs_per_s"] = M * K / runs["fp64_plain"]["secs"]
        out["fp64_fused_moment_vecs_per_s"] = M * K / runs["fp64_fused"]["secs"]
        out["speedup_fp32_mixed_vs_fp64_plain"] = runs["fp64_plain"]["secs"] / fp32["secs"]
        out["dos_error_fp32_vs_fp64"] = dos_err
        out["dos_error_fp64_fused_vs_fp64_plain"] = KP.dos_error(rho64f, rho64)
        out["moments_per_matvec"] = {"fp64_plain": 1, "fp64_fused": 2, "fp32_mixed": 2}
        # roofline of the dominant kernel: the fused FP32 step (k_apply_lanes, 4x4 blocks, 10 columns)
        per_launch_ms = fp32["kernel_ms"] / max(fp32["launches"], 1)
        alg = spmm_bytes(G, K, 4, 4, with_z=True)
        achieved = alg / (per_launch_ms * 1e-3) / 1e9
        roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": achieved / HBM_PEAK_GBS,
                    "traffic": read_traffic("pmc_kpm_step_fp32_graphene200.txt"),
                    "kernel": "k_apply_lanes<float,float,4,1,10,2,true,0> (one fused KPM step)",
                    "alg_bytes_per_launch": alg, "avg_launch_us": per_launch_ms * 1e3,
                    "note": "working set ~95 MB fits the 256 MB Infinity Cache; see spmv_large for HBM"}
        p64 = runs["fp64_fused"]
        pl64 = p64["kernel_ms"] / max(p64["launches"], 1)
        a64 = spmm_bytes(G, K, 8, 8, with_z=True)
        out["fp64_fused_step_gbs"] = a64 / (pl64 * 1e-3) / 1e9
        if not args.quick:
            out["kpm_hbm_resident"] = bench_kpm_scaled(args, sync, log)

    if want("spmv"):
        # block SpMV FP64 on a 1.5 GB variable-block matrix (>> 256 MB MALL)
        nb = 20000 if args.quick else 130000
        log(f"generating random VBCSR nb={nb} for SpMV bandwidth")
        B = W.random_vbcsr(nb, per_row=40, seed=1)
        HB = VBCSR.from_block_csr(B.bs, B.row_ptr, B.col_ind, B.vals)
        x = torch.rand(B.n, dtype=torch.float64, device="cuda")
        y = torch.zeros(B.n, dtype=torch.float64, device="cuda")
        core = HB._core
        core.spmm_raw(1, x.data_ptr(), 1, y.data_ptr(), 1, 0, False)
        sync()
        reps = max(args.steps, 20)
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ev0.record()
        for _ in range(reps):
            core.spmm_raw(1, x.data_ptr(), 1, y.data_ptr(), 1, 0, False)
        ev1.record()
        sync()
        ms = ev0.elapsed_time(ev1) / reps
        bts = 8 * B.nnz + 4 * B.nnzb + 4 * (B.nb + 1) + 8 * B.n * 2
        gbs = bts / (ms * 1e-3) / 1e9
        out["spmv_large"] = {"matrix": f"random VBCSR nb={nb}, blocks 4-8, diag + 40 random cols/row, "
                                       f"{B.nnz * 8 / 1e9:.2f} GB FP64 values",
                             "ms": ms, "gbs": gbs, "frac_hbm": gbs / HBM_PEAK_GBS, "alg_bytes": bts}
        log(f"spmv large: {ms:.3f} ms, {gbs:.1f} GB/s ({100 * gbs / HBM_PEAK_GBS:.1f}% of HBM peak)")
        # the box's streaming ceiling on the same byte count: a plain read
        # (torch sum over FP64) and a copy (half read, half written)
        buf = torch.empty(bts // 8, dtype=torch.float64, device="cuda").uniform_()
        dst = torch.empty(bts // 16, dtype=torch.float64, device="cuda")
        half = buf[: bts // 16]
        ceil = {}
        for name, fn in (("read", lambda: buf.sum()), ("copy", lambda: dst.copy_(half))):
            fn()
            sync()
            ev0.record()
            for _ in range(reps):
                fn()
            ev1.record()
            sync()
            t = ev0.elapsed_time(ev1) / reps * 1e-3
            ceil[f"{name}_gbs"] = (8 * (bts // 8) if name == "read" else 16 * (bts // 16)) / t / 1e9
        del buf, dst, half
        ceil["spmv_frac_of_read"] = gbs / ceil["read_gbs"]
        out["spmv_large"]["stream_ceiling"] = ceil
        log(f"streaming ceiling on {bts / 1e9:.2f} GB: read {ceil['read_gbs']:.0f} GB/s, copy {ceil['copy_gbs']:.0f} "
            f"GB/s; SpMV at {100 * ceil['spmv_frac_of_read']:.0f}% of the read rate")
        if not args.no_cpu and not args.quick:
            import oracle

            if oracle.ref_available():
                # the reference's BlockSpMat::mult on the same matrix and x, OpenMP threads
                R = oracle.RefMat(oracle.Mat.of(B))
                xs = x.cpu().numpy()
                cpu_s = R.time_mult(xs, 5)
                out["spmv_large"]["cpu_reference"] = {
                    "ms": cpu_s * 1e3, "gbs": bts / cpu_s / 1e9, "gpu_speedup": cpu_s * 1e3 / ms,
                    "threads": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)),
                    "kind": "reference", "sample": "vbcsr C++ core BlockSpMat::mult (oracle/_ref), mean of 5 calls "
                                                   "after a warm-up"}
                log(f"spmv large, reference CPU: {cpu_s * 1e3:.1f} ms")
                del R
        # the same on the graphene matrix, FP64, K=1 (cache-resident)
        xg = torch.rand(G.n, dtype=torch.float64, device="cuda")
        yg = torch.zeros(G.n, dtype=torch.float64, device="cuda")
        ev0.record()
        for _ in range(reps):
            H._core.spmm_raw(1, xg.data_ptr(), 1, yg.data_ptr(), 1, 0, False)
        ev1.record()
        sync()
        msg = ev0.elapsed_time(ev1) / reps
        bg = 8 * G.nnz + 4 * G.nnzb + 4 * (G.nb + 1) + 8 * G.n * 2
        out["spmv_graphene"] = {"ms": msg, "gbs": bg / (msg * 1e-3) / 1e9}
        del HB, x, y

    if want("batched"):
        count = 16 if args.quick else 64
        Mb, Kb = 512, 32
        log(f"generating {count} batched Hamiltonians")
        Hs = W.batched_hamiltonians(count=count, seed=7, scale=False)
        mats = [VBCSR.from_block_csr(A.bs, A.row_ptr, A.col_ind, A.vals) for A in Hs]
        # each spectrum mapped onto [-0.99, 0.99] by its own device-Lanczos bounds
        t_sc = time.perf_counter()
        bounds = [KP.scale_to_unit(Hm, 0.99) for Hm in mats]
        sync()
        log(f"spectral scaling of {count} Hamiltonians (device Lanczos, 120 steps): "
            f"{time.perf_counter() - t_sc:.2f} s")
        V0s = []
        for h, Hm in enumerate(mats):
            v = Hm.create_multivector(Kb)
            v.set_random(1000 + h, normalize=False)
            V0s.append(v)
        sync()
        res = {}
        legs = [g for g in args.batched_legs.split(",") if g]
        if args.batched_plan:
            KP.set_batch_plan(*map(int, args.batched_plan.split(":")))

        def batched(prec):
            def run():
                res[prec] = KP.kpm_moments_batched(mats, V0s, Mb, mode="two", precision=prec, graph=True)
            return run

        def plain_fp64():
            res["fp64"] = np.stack([KP.kpm_moments(Hm, v, Mb, mode="plain") for Hm, v in zip(mats, V0s)])

        bsteps = max(2, args.steps // 5)
        tb, plans, step_us = {}, {}, {}
        precs = [g for g in ("mixed", "fp32") if g in legs]
        for prec in precs:
            tb[prec] = timed(batched(prec), bsteps, 1, sync)
            t_, p_, tuned_ = KP.last_plan()
            plans[prec] = {"class_threshold": t_, "slot_parts": p_, "autotuned": tuned_}
            kms, steps_ = vbcsr_amd._core.kpm_last_step_time()  # the graph's time over its recursion steps
            step_us[prec] = kms * 1e3 / max(steps_, 1)
        if args.batched_plan:
            KP.set_batch_plan(0)
        rows = int(sum(A.n for A in Hs))
        bres = {"workload": f"{count} random-geometric Hamiltonians, blocks {{1,4,9,13}}, "
                            f"{rows} rows total, each spectrum scaled to [-0.99, 0.99] by its "
                            f"device-Lanczos bounds, {Mb} moments x {Kb} vectors each",
                "spectral_bounds_first4": [list(b) for b in bounds[:4]],
                "nnz_total": int(sum(A.nnz for A in Hs)),
                "plan_note": "class_threshold < 0: class rows on the block-row wave kernel (vector ALUs), > 0: "
                             "on the FP64 matrix cores; slot_parts < 0: row-wave kernel rows per work-group"}
        t64 = timed(plain_fp64, max(1, bsteps // 2), 1, sync) if "fp64" in legs else None
        if t64 is not None:
            bres["fp64_plain_moment_vecs_per_s"] = count * Mb * Kb / t64
        for prec in precs:
            r = {"moment_vecs_per_s": count * Mb * Kb / tb[prec], "us_per_step": step_us[prec], "plan": plans[prec]}
            if t64 is not None:
                e = []
                for h in range(count):
                    _, r64 = KP.dos_from_moments(res["fp64"][h])
                    _, rp = KP.dos_from_moments(res[prec][h])
                    e.append(KP.dos_error(rp, r64))
                r["speedup_vs_fp64_plain"] = t64 / tb[prec]
                r["max_dos_error"] = float(max(e))
            bres[prec] = r
            log(f"batched {prec}: plan {plans[prec]}, {tb[prec] * 1e3:.2f} ms ({step_us[prec]:.1f} us per step)"
                + (f", speedup {r['speedup_vs_fp64_plain']:.2f}x vs fp64 plain {t64 * 1e3:.2f} ms, "
                   f"max DOS err {r['max_dos_error']:.2e}" if t64 is not None else ""))
        if not args.no_cpu and not args.quick:
            import oracle

            if oracle.ref_available():
                # the plain FP64 ex08 composition of the reference core on every
                # 8th Hamiltonian (device-scaled values, same vectors), scaled
                # to the batch by stored values
                from vbcsr_amd.workloads import BlockCSR

                secs, nnz_s = 0.0, 0
                for h in range(0, count, 8):
                    A = Hs[h]
                    vals = np.asarray(mats[h].get_values())
                    R = oracle.RefMat(oracle.Mat.of(BlockCSR(A.bs, A.row_ptr, A.col_ind, vals)))
                    _, t_ = R.kpm_moments(np.asarray(V0s[h].to_numpy()), Mb)
                    secs += t_
                    nnz_s += A.nnz
                est = secs * bres["nnz_total"] / nnz_s
                bres["cpu_reference"] = {
                    "sample_hamiltonians": len(range(0, count, 8)), "sample_seconds": secs,
                    "batch_seconds_estimate": est,
                    "threads": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)), "kind": "reference",
                    "sample": f"vbcsr C++ core (oracle/_ref), ex08 composition, FP64, {Mb} moments x {Kb} vectors on "
                              "every 8th Hamiltonian, scaled to the batch by stored values",
                    "gpu_mixed_speedup": est / tb["mixed"] if "mixed" in tb else None}
                log(f"batched, reference CPU: {secs:.2f} s on the sample -> {est:.1f} s for the batch")
        out["batched"] = bres

    if want("prop"):
        out["propagation"] = bench_propagation(args, sync, log)

    if want("mcweeny"):
        out["mcweeny"] = bench_mcweeny(args, sync, log)

    cpu = None
    if want("cpu") and not args.no_cpu and not args.quick:
        log("CPU baseline (reference core on the host)")
        try:
            cpu, _ = cpu_baseline(G, K, M, seed)
            cpu["host"] = host_info()
            log(f"cpu: {cpu['value']:.4g} moment-vecs/s ({cpu['kind']}, {cpu['cores']} threads, "
                f"{cpu['host'].get('Model name', '?')})")
        except Exception as e:  # the baseline leg must not hide the GPU result
            cpu = {"value": None, "unit": "moment-vecs/s", "cores": None, "kind": "reference",
                   "sample": f"failed: {e}"}

    line = {
        "metric": "KPM moment-vecs/s + block-SpMV GB/s, FP32-mixed vs FP64 at stated DOS error",
        "value": out.get("value"),
        "unit": "moment-vecs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (FP64 moment reductions)",
        "data": "synthetic (seeded graphene-like Hamiltonian, Rademacher vectors)",
        "config": {"workload": f"graphene {nx}x{nx} cells, 4 orbitals (N={G.n}), KPM {M} moments, "
                               f"{K} Rademacher vectors, Jackson kernel",
                   "global_batch": K, "seq_len": M, "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    line.update({k: v for k, v in out.items() if k != "value"})
    print(json.dumps(line))


if __name__ == "__main__":
    main()
