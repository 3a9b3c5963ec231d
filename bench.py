"""MPO.MPS compress sweeps/s on one MI355X.

Workload (the configuration the headline metric is quoted on): XXZ MPO
(D=5, Jz=1, Jxy=1, hz=0.5) applied to a seeded random MPS (N=64, d=2,
bonds [1,512,...,512,1], i.i.d. N(0,1) FP64 entries, normalised) and
compressed back to chi=512 (threshold 1e-15).  One step = one sweep =
MPO application + compression of the same resident input state.

  value  -- sweeps/s of the fast path: the structured apply+compress sweep
            (tncontract_amd.onedim.structured) with mixed-precision
            contractions, FP64 Jacobi rotations and FP64 output isometries.
  plain  -- the reference algorithm in FP64 on the GPU (contract_mps_mpo +
            svd_compress: QR left sweep on the chi*D chain, truncating SVD
            right sweep), timed in the same run for the ratio.
  err    -- ||psi_fast - psi_plain||^2 / ||psi_plain||^2 (FP64 overlaps on the
            GPU); the plain path itself is checked against the CPU oracle by
            the test suite.
  roofline -- the dominant kernel of the fast sweep, timed live with HIP
            events on its launch stream (library launch timers).
  cpu_baseline -- the NumPy FP64 oracle's full reference sweep of a 22-site
            chain of the workload (a bounded sample), scaled to the 64-site
            sweep by the per-site flop model.
  oracle_check -- the fast and plain GPU paths on that 22-site chain
            against the oracle's result (state error, bond dimensions).

Prints ONE JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("MPO·MPS compress sweeps/s at χ=512: mixed+learned vs plain FP64 GPU "
          "@1e-8 err")
PEAK_TFLOPS = {"gemm_f64": 78.6, "gemm_f32": 157.3, "gemm_bf16x3": 2500.0,
               "jac_gram": 78.6, "jac_eig_apply": 78.6, "jac_apply": 78.6}
NOTES = {"jac_eig_apply": "pair eigensolve (FP64 rotations, 31 dependent rounds with one barrier "
                          "each) + MFMA apply; latency-bound"}
# HBM bytes per launch (FETCH_SIZE x 2.000 + WRITE_SIZE x 1.000, calibrated on a
# streaming kernel) measured for the chi=512 sweep's n=1024 Jacobi launches;
# profiles/jac_eig_chi512_pmc_traffic.md.  Used when --traffic is not given.
MEASURED_TRAFFIC = {"jac_eig_apply": 16738191, "jac_gram": 10560512}


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--N", type=int, default=64)
    p.add_argument("--chi", type=int, default=512)
    p.add_argument("--seed", type=int, default=0)
    p.add_argument("--precision", default="fp32", choices=["fp64", "fp32", "bf16x3"])
    p.add_argument("--plain-steps", type=int, default=1)
    p.add_argument("--no-plain", action="store_true")
    p.add_argument("--eig-tol", type=float, default=None,
                   help="Jacobi tolerance of the fast path's eigensolver (default: the "
                        "structured sweep's own, 1e-8 for mixed precision)")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline sample")
    p.add_argument("--no-extra", action="store_true",
                   help="headline only: skip the config 3 and config 4 runs reported in its line")
    p.add_argument("--traffic", type=float, default=None,
                   help="HBM bytes per launch of the dominant kernel, from the calibrated "
                        "FETCH_SIZE + WRITE_SIZE counters (reported as roofline.traffic)")
    p.add_argument("--workload", default="chain",
                   choices=["chain", "batched", "variational", "peps", "tebd", "learned",
                            "config5", "kernels"],
                   help="chain: the MPO.MPS compress sweep (default, the headline); batched: "
                        "1024 uneven contraction + truncation jobs (chi ~ U{16..512}, d=2, "
                        "FP32 data, keep sigma/sigma_0 > 1e-8); variational: iterations/s of "
                        "variational_compress on the MPO-applied chain; peps: boundary-MPS "
                        "inner_product_peps contractions/s; tebd: canonical-MPS two-site "
                        "gates/s; learned: learned-rank compression sweeps/s (config 4); "
                        "config5: chi=1024 bf16x3 sweeps, 4 in a row; kernels: the hot "
                        "kernels at the headline shapes, for counter profiles")
    p.add_argument("--tebd-N", type=int, default=64)
    p.add_argument("--tebd-chi", type=int, default=128)
    p.add_argument("--peps-L", type=int, default=10)
    p.add_argument("--peps-D", type=int, default=3)
    p.add_argument("--peps-chi", type=int, default=64)
    p.add_argument("--jobs", type=int, default=1024)
    p.add_argument("--copy-calibration", type=int, default=0, metavar="MB",
                   help="only run a plain device copy of MB megabytes three times (the "
                        "reference kernel for scaling the FETCH_SIZE / WRITE_SIZE counters) "
                        "and exit")
    return p.parse_args()


def copy_calibration(mb):
    """Three launches of the library's streaming kernel tn_scale (x *= 1 on an
    FP64 buffer of mb MB): exactly mb MB read and mb MB written per launch."""
    import torch
    from tncontract_amd import _lib as L
    torch.cuda.set_device(0)
    n = mb * (1 << 20) // 8
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    for _ in range(3):
        L.scale_(x, 1.0)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_read": n * 8, "copy_calibration_bytes_written": n * 8,
                      "launches": 3, "check": float(x[0])}), flush=True)


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def flops_per_site_plain(chi_l, chi_r, D_l, D_r, d, chi_out_r, qb_l):
    """Reference algorithm work for one site (closed forms): QR of the
    (d*qb_l) x (chi_r D_r) fused site, R . next-site GEMM, truncating SVD."""
    m, n = d * qb_l, chi_r * D_r
    k = min(m, n)
    qr = 4.0 * (m * n * k - (m + n) * k * k / 2.0 + k ** 3 / 3.0) + (2.0 * m * k * k - 2.0 * k ** 3 / 3.0)
    gemm = 2.0 * k * n * (d * chi_r * D_r)
    p, q = max(d * chi_out_r, qb_l), min(d * chi_out_r, qb_l)
    svd = 6.0 * p * q * q + 20.0 * q ** 3
    return qr + gemm + svd


def sweep_site_equivalents(N, d, chi, D=5):
    """Bulk-site equivalents of the reference sweep on an N-site chain with
    bonds [1, chi, ..., chi, 1]: the per-site flop model summed over the
    sites, with the QR sweep's left bonds min(d * previous, chi D) and the
    truncated right bonds min(chi, d^(N-1-i)), divided by a bulk site's."""
    qb = [1]
    for i in range(N - 1):
        qb.append(min(d * qb[-1], chi * D))
    bulk = flops_per_site_plain(chi, chi, D, D, d, chi, chi * D)
    return sum(flops_per_site_plain(chi, chi, D, D, d, min(chi, d ** (N - 1 - i)), qb[i]) / bulk
               for i in range(1, N - 1))


def cpu_baseline_sweep(args, N_small=22, d=2, D=5):
    """The oracle's full reference sweep (contract_mps_mpo, then
    svd_compress(chi, 1e-15)) of an N_small-site chain at the workload's chi,
    timed on the host cores and scaled to the N-site sweep by the per-site
    flop model; its result also checks the GPU paths on the same chain
    (state error ||psi - psi_oracle||^2 / ||psi_oracle||^2, bond dimensions).
    Returns (cpu_baseline record, oracle check record)."""
    import numpy as np
    import torch
    import tncontract_amd as tn
    from tncontract_amd.onedim import random_mps_arrays, xxz_mpo
    from tncontract_amd.onedim.structured import apply_mpo_compress, SweepStats
    from oracle import tn_oracle as o
    from oracle import models as om
    try:
        from threadpoolctl import threadpool_info
        blas = [(x.get("internal_api"), x.get("num_threads")) for x in threadpool_info()]
    except Exception:  # pragma: no cover
        blas = []
    cores = len(os.sched_getaffinity(0))
    chi = args.chi
    arrs = random_mps_arrays(N_small, d, chi, seed=args.seed)
    t0 = time.perf_counter()
    r = o.contract_mps_mpo(o.OMPS([o.OTensor(a, ["phys", "left", "right"]) for a in arrs]),
                           om.xxz_mpo(N_small))
    t_mpo = time.perf_counter() - t0
    r.svd_compress(chi=chi, threshold=1e-15)
    t_all = time.perf_counter() - t0
    scale = sweep_site_equivalents(args.N, d, chi, D) / sweep_site_equivalents(N_small, d, chi, D)
    cpu = {"value": 1.0 / (t_all * scale), "unit": "sweeps/s", "cores": cores, "kind": "port",
           "sample": ("NumPy FP64 oracle, the full reference sweep (contract_mps_mpo %.2fs, then "
                      "svd_compress(chi=%d, 1e-15): QR left sweep + truncating SVD right sweep) "
                      "of a %d-site chain of the workload (%.2fs), scaled to the %d-site sweep "
                      "by the closed-form per-site flop model (x%.2f); BLAS %s"
                      % (t_mpo, chi, N_small, t_all, args.N, scale, blas))}

    def to_oracle(g):
        return o.OMPS([o.OTensor(t.numpy().astype(np.float64), t.labels) for t in g],
                      g.left_label, g.right_label, g.phys_label)

    def err(g):
        a = to_oracle(g)
        nb = o.inner_product_mps(r, r)
        return float((o.inner_product_mps(a, a) + nb - 2 * o.inner_product_mps(a, r)) / nb)

    psi = tn.onedim.MatrixProductState([tn.Tensor(a, ["phys", "left", "right"]) for a in arrs])
    st = SweepStats()
    fast = apply_mpo_compress(psi, xxz_mpo(N_small), chi=chi, threshold=1e-15,
                              precision=args.precision, stats=st, eig_tol=args.eig_tol)
    plain = tn.onedim.contract_mps_mpo(psi, xxz_mpo(N_small))
    plain.svd_compress(chi=chi, threshold=1e-15)
    torch.cuda.synchronize()
    check = {"chain": "XXZ MPO D=5 x random MPS N=%d d=%d chi=%d -> chi=%d (seed %d)"
                      % (N_small, d, chi, chi, args.seed),
             "fast_state_err": err(fast), "fast_bonddims_equal": fast.bonddims() == r.bonddims(),
             "fast_fallback_bond": st.fallback,
             "plain_fp64_state_err": err(plain),
             "plain_bonddims_equal": plain.bonddims() == r.bonddims()}
    return cpu, check


# FP64-equivalent peaks (TFLOP/s) of the tagged kernels for the sweep-level
# roofline; bf16x3 takes three BF16 MFMAs per FP32-equivalent product
SWEEP_PEAKS = {"gemm_f64": 78.6, "gemm_f32": 157.3, "gemm_bf16x3": 2500.0 / 3.0,
               "jac_gram": 78.6, "jac_eig_apply": 78.6, "jac_apply": 78.6, "qr_panel": 78.6,
               "direct": None}


def sweep_roofline(prof, t_measured_s):
    """T_min = sum over the tagged kernels of max(F_k / P_k, B_k / 8 TB/s)
    (F_k, B_k: the algorithmic flops and bytes the library counted for them
    in one profiled sweep) against the measured sweep time."""
    t_min = 0.0
    for tag, (ms, n, fl, by) in prof.items():
        if not n:
            continue
        peak = SWEEP_PEAKS.get(tag)
        tf = fl / (peak * 1e12) if peak else 0.0
        t_min += max(tf, by / 8e12)
    return {"t_min_ms": t_min * 1e3, "t_measured_ms": t_measured_s * 1e3,
            "frac": t_min / t_measured_s if t_measured_s > 0 else None,
            "note": "T_min over the kernels the library times (GEMMs, Jacobi, QR panels, "
                    "thin-K contractions) at their closed-form work"}


def batched_workload(args):
    """Jobs/s of the batched uneven contraction + truncation (all jobs resident
    on the device; one step = every job's GEMM + truncated SVD)."""
    import numpy as np
    import torch
    from tncontract_amd import _lib as L
    from tncontract_amd.batched import batched_truncate, config3_jobs
    torch.cuda.set_device(0)
    jobs = config3_jobs(args.jobs, seed=args.seed)
    Ls = [torch.from_numpy(j[0]).cuda() for j in jobs]
    Rs = [torch.from_numpy(j[1]).cuda() for j in jobs]
    torch.cuda.synchronize()
    log("batched: %d jobs warmup" % len(jobs))
    for _ in range(args.warmup):
        batched_truncate(Ls, Rs, rel_threshold=1e-8)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        U, S, Vt, kept, sweeps = batched_truncate(Ls, Rs, rel_threshold=1e-8)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    # parity sample: kept rank and singular values of 8 jobs against LAPACK
    # on the same data (the oracle's FP64 product of the FP32 factors)
    err, kept_ok = 0.0, True
    for j in range(0, len(jobs), max(1, len(jobs) // 8)):
        C = jobs[j][0].astype(np.float64) @ jobs[j][1].astype(np.float64)
        ref = np.linalg.svd(C, compute_uv=False)
        k = int(kept[j])
        kept_ok = kept_ok and k == int(np.count_nonzero(ref / ref[0] > 1e-8))
        err = max(err, float(np.max(np.abs(S[j].cpu().numpy()[:k] - ref[:k])) / ref[0]))
    cpu = None
    if not args.no_cpu:
        from oracle import tn_oracle as o
        # the whole batch (about 20 s on 16 host cores), capped at 60 s of work
        t1 = time.perf_counter()
        done = 0
        for Lm, Rm, _ in jobs:
            A = o.OTensor(Lm.astype(np.float64), ["l", "m"])
            B = o.OTensor(Rm.astype(np.float64), ["m", "r"])
            o.truncated_svd(o.contract(A, B, "m", "m"), ["l"], threshold=1e-8, absolute=False)
            done += 1
            if time.perf_counter() - t1 > 60.0:
                break
        tt = time.perf_counter() - t1
        cpu = {"value": done / tt, "unit": "jobs/s", "cores": len(os.sched_getaffinity(0)),
               "kind": "port", "sample": "NumPy FP64 oracle contract + truncated_svd of %d of the "
                                         "%d jobs, in order (%.2f s)" % (done, len(jobs), tt)}
    L.prof_collect("gemm_f64", reset=True)
    L.prof_enable(True)
    batched_truncate(Ls, Rs, rel_threshold=1e-8)
    torch.cuda.synchronize()
    L.prof_enable(False)
    prof = {tag: L.prof_collect(tag) for tag in L.PROF_TAGS}
    L.prof_collect("gemm_f64", reset=True)
    roof = roofline_from_prof(prof, args.traffic)
    roof["note"] = ("batched Jacobi work is counted on the device: rows and eigensolves of "
                    "the jobs and block pairs still rotating")
    flops = sum(2.0 * j[0].shape[0] * j[0].shape[1] * j[1].shape[1] for j in jobs)
    rec = {"metric": "batched uneven contraction + truncation jobs/s", "value": len(jobs) / dt,
           "unit": "jobs/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": dt * 1e3, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64 (FP32 factors)",
           "data": "synthetic (seeded factors with exp(-k/tau) spectra)",
           "config": {"workload": "%d jobs, chi ~ U{16..512}, d=2, keep sigma/sigma_0 > 1e-8"
                                  % len(jobs)},
           "jacobi_sweeps_max": int(np.max(sweeps)), "kept_mean": float(np.mean(kept)),
           "kept_equal_lapack_sample": kept_ok,
           "gemm_gflop": flops / 1e9, "sv_rel_err_sample": err,
           "kernel_ms_profiled_step": {t: round(v[0], 3) for t, v in prof.items() if v[1]},
           "roofline": roof, "cpu_baseline": cpu}
    return rec


def roofline_from_prof(prof, traffic=None, measured=None):
    """Roofline record for the tag with the largest measured time; `measured`
    maps kernel tags to PMC-measured HBM bytes per launch for this workload."""
    measured = measured or {}
    dom = max(prof, key=lambda t: prof[t][0])
    ms, n, fl, by = prof[dom]
    # the binding roof by arithmetic intensity: below the ridge point
    # (peak FLOP/s / 8 TB/s) HBM bounds the kernel, above it MFMA does
    mfma_bound = dom in PEAK_TFLOPS and fl > 0 and (
        by <= 0 or fl / by >= PEAK_TFLOPS[dom] * 1e12 / 8.0e12)
    if mfma_bound:
        ach = fl / (ms * 1e-3) / 1e12
        roof = {"bound": "mfma", "kernel": dom, "achieved": ach, "peak": PEAK_TFLOPS[dom],
                "unit": "TFLOP/s", "frac": ach / PEAK_TFLOPS[dom],
                "traffic": traffic if traffic is not None else measured.get(dom),
                "avg_launch_us": ms * 1e3 / max(n, 1), "launches": n,
                "flops_per_launch": fl / max(n, 1)}
    else:
        ach = by / (ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "kernel": dom, "achieved": ach, "peak": 8000.0, "unit": "GB/s",
                "frac": ach / 8000.0,
                "traffic": traffic if traffic is not None else measured.get(dom),
                "avg_launch_us": ms * 1e3 / max(n, 1), "launches": n,
                "bytes_per_launch": by / max(n, 1)}
        if fl > 0:
            roof["flops_per_launch"] = fl / max(n, 1)
            roof["flop_per_byte"] = fl / by
    if dom in NOTES:
        roof["note"] = NOTES[dom]
    return roof


def variational_cpu_sample(chi, chiD, d=2):
    """The oracle's bulk site update of variational_compress (onedim_core.py:
    548-597) at the workload's shapes: target . left env, . right env, LQ,
    L into the neighbour, right-environment update."""
    import numpy as np
    from oracle import tn_oracle as o
    rng = np.random.default_rng(2)
    tgt = o.OTensor(rng.standard_normal((chiD, d, chiD)), ["left", "phys", "right"])
    le = o.OTensor(rng.standard_normal((chi, chiD)), ["le1", "le2"])
    re = o.OTensor(rng.standard_normal((chi, chiD)), ["re1", "re2"])
    nb = o.OTensor(rng.standard_normal((chi, d, chi)), ["l", "p", "r"])
    t0 = time.perf_counter()
    upd = o.contract(tgt, le, "left", "le2")
    upd = o.contract(upd, re, "right", "re2")
    Lt, Q = o.tensor_lq(upd, "le1", lq_label="lq")
    o.contract(nb, Lt, "r", "le1")
    r2 = o.contract(Q, re, "re1", "re1")
    o.contract(r2, tgt, ["phys", "re2"], ["phys", "right"])
    return time.perf_counter() - t0


def variational_workload(args):
    """Iterations/s of MatrixProductState.variational_compress on the XXZ MPO
    applied to the workload MPS (target bond chi*D), compressed to chi, with
    the structured sweep's output as the initial guess.  One step = one full
    iteration of the reference loop (a right-to-left and a mirrored sweep)."""
    import numpy as np
    import torch
    import tncontract_amd as tn
    from tncontract_amd import _lib as L
    from tncontract_amd.onedim import random_mps_arrays, xxz_mpo
    from tncontract_amd.onedim.onedim_core import _variational_compress
    from tncontract_amd.onedim.structured import apply_mpo_compress
    torch.cuda.set_device(0)
    N, d, chi, D = args.N, 2, args.chi, 5
    arrs = random_mps_arrays(N, d, chi, seed=args.seed)
    psi = tn.onedim.MatrixProductState([tn.Tensor(a, ["phys", "left", "right"]) for a in arrs])
    mpo = xxz_mpo(N)
    target = tn.onedim.contract_mps_mpo(psi, mpo)
    guess = apply_mpo_compress(psi, mpo, chi=chi, threshold=1e-15, precision="fp64")
    torch.cuda.synchronize()
    log("variational: target bonds %d, guess bonds %d" % (max(target.bonddims()), max(guess.bonddims())))

    def run(iters, prof=False):
        g = guess.copy()
        torch.cuda.synchronize()
        if prof:
            L.prof_collect("gemm_f64", reset=True)
            L.prof_enable(True)
        t0 = time.perf_counter()
        out, it, _ = _variational_compress(target, chi, iters, g, 0.0, False)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        if prof:
            L.prof_enable(False)
        assert it == iters
        return out, dt

    log("variational warmup")
    run(1)
    # one-iteration run gives the fixed set-up cost (QR left sweep of the
    # guess + left environments) plus one iteration; the timed run adds steps
    _, t1 = run(1)
    log("variational timed x%d" % args.steps)
    out, tk = run(1 + args.steps)
    dt = (tk - t1) / args.steps
    # projection identities of the final single-site update (exact at any
    # iteration count): <w|v> = <w|w>, |v - w|^2 = |v|^2 - |w|^2
    ip = tn.onedim.inner_product_mps
    vv, ww, wv = float(ip(target, target)), float(ip(out, out)), float(ip(out, target))
    proj_err = abs(wv - ww) / ww
    dist_var = (vv - ww) / vv
    dist_guess = float(tn.onedim.frob_distance_squared(guess, target)) / vv
    L.prof_collect("gemm_f64", reset=True)
    _, _ = run(1, prof=True)
    prof = {tag: L.prof_collect(tag) for tag in L.PROF_TAGS}
    L.prof_collect("gemm_f64", reset=True)
    roof = roofline_from_prof(prof, args.traffic)
    cpu = None
    if not args.no_cpu:
        log("cpu baseline sample")
        ts = variational_cpu_sample(chi, chi * D, d)
        # an iteration visits 2(N-1) site updates; edge sites are smaller,
        # counted as bulk (the sample overstates the CPU rate slightly)
        per_iter = ts * 2 * (N - 1)
        cpu = {"value": 1.0 / per_iter, "unit": "iterations/s", "cores": len(os.sched_getaffinity(0)),
               "kind": "port", "sample": "NumPy FP64 oracle, one bulk site update of "
               "variational_compress at chi=%d, target bond %d (%.2f s), times the %d site "
               "updates of an iteration" % (chi, chi * D, ts, 2 * (N - 1))}
    rec = {"metric": "variational_compress iterations/s", "value": 1.0 / dt,
           "unit": "iterations/s", "n_gpus": 1, "steps": args.steps, "warmup": 1,
           "ms_per_step": dt * 1e3, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64",
           "data": "synthetic (seeded random MPS, i.i.d. N(0,1), normalised; XXZ MPO)",
           "config": {"workload": "variational_compress of XXZ MPO D=5 x random MPS N=%d d=%d "
                                  "chi=%d (target bond %d) to chi=%d, structured-sweep initial guess"
                                  % (N, d, chi, chi * D, chi)},
           "setup_plus_one_iteration_s": t1,
           "rel_dist2_guess": dist_guess, "rel_dist2_variational": dist_var,
           "projection_identity_rel_err": proj_err,
           "kernel_ms_profiled_iteration": {t: round(v[0], 3) for t, v in prof.items() if v[1]},
           "roofline": roof, "cpu_baseline": cpu}
    return rec


def peps_arrays(L, D, d, seed):
    """Seeded random L x L PEPS site arrays (phys, up, right, down, left),
    i.i.d. N(0,1) scaled by 1/sqrt(d D^2) so the network stays O(1)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    out = []
    for i in range(L):
        row = []
        for j in range(L):
            shape = (d, 1 if i == 0 else D, 1 if j == L - 1 else D, 1 if i == L - 1 else D,
                     1 if j == 0 else D)
            row.append(rng.standard_normal(shape) / np.sqrt(d * D * D))
        out.append(row)
    return out


def peps_workload(args):
    """Contractions/s of inner_product_peps(psi, psi) by boundary MPS
    (square_lattice.py:136-203, 281-304): L x L PEPS of bond D, double-layer
    bond D^2, boundary MPS compressed to chi by svd_compress_mps per column.
    One step = one full contraction (all columns)."""
    import numpy as np
    import torch
    import tncontract_amd as tn
    from tncontract_amd import _lib as L
    torch.cuda.set_device(0)
    Ls, D, d, chi = args.peps_L, args.peps_D, 2, args.peps_chi
    labels = ["phys", "up", "right", "down", "left"]
    arrs = peps_arrays(Ls, D, d, args.seed)
    psi = tn.twodim.SquareLatticePEPS([[tn.Tensor(a, list(labels)) for a in row] for row in arrs])
    torch.cuda.synchronize()

    def step():
        return tn.twodim.inner_product_peps(psi, psi, exact_contract=False, chi=chi)

    log("peps: L=%d D=%d chi=%d warmup" % (Ls, D, chi))
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    log("peps timed x%d" % args.steps)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        val = step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    value = float(val.numpy().reshape(-1)[0])
    L.prof_collect("gemm_f64", reset=True)
    L.prof_enable(True)
    step()
    torch.cuda.synchronize()
    L.prof_enable(False)
    prof = {tag: L.prof_collect(tag) for tag in L.PROF_TAGS}
    L.prof_collect("gemm_f64", reset=True)
    roof = roofline_from_prof(prof, args.traffic)
    cpu, rel_err = None, None
    if not args.no_cpu:
        from oracle import tn_oracle as o
        from oracle import square_lattice_oracle as sq
        log("cpu baseline: oracle on the same network")
        ops = sq.OPEPS([[o.OTensor(a, list(labels)) for a in row] for row in arrs])
        t1 = time.perf_counter()
        ref = sq.inner_product_peps(ops, ops, exact_contract=False, chi=chi)
        tc = time.perf_counter() - t1
        refv = float(np.asarray(ref.data, dtype=np.float64).reshape(-1)[0])
        rel_err = abs(value - refv) / abs(refv)
        cpu = {"value": 1.0 / tc, "unit": "contractions/s", "cores": len(os.sched_getaffinity(0)),
               "kind": "port", "sample": "NumPy FP64 oracle, the same full contraction once "
                                         "(%.2f s)" % tc}
    rec = {"metric": "inner_product_peps boundary-MPS contractions/s", "value": 1.0 / dt,
           "unit": "contractions/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": dt * 1e3, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64",
           "data": "synthetic (seeded random PEPS, i.i.d. N(0,1)/sqrt(d D^2))",
           "config": {"workload": "inner_product_peps(psi, psi), %dx%d PEPS d=%d D=%d "
                                  "(double-layer bond %d), mps_contract svd chi=%d"
                                  % (Ls, Ls, d, D, D * D, chi)},
           "value_norm2": value, "rel_err_vs_oracle": rel_err,
           "kernel_ms_profiled_step": {t: round(v[0], 3) for t, v in prof.items() if v[1]},
           "roofline": roof, "cpu_baseline": cpu}
    return rec


def kernels_workload(args):
    """The headline sweep's hot kernels at their shapes, each launched a few
    times on resident data (the driver for the PMC counter profiles):
      gemm_f64 / gemm_f32 / gemm_bf16x3: the R . next-site contraction of the
        plain sweep, (2560 x 2560) . (2560 x 2 x 2560), FP64 / FP32 / bf16x3;
      permute: a (2560, 2, 2560) FP64 site moved to (2, 2560, 2560) order;
      direct: the MPO application of one site, (2, 512, 512) x XXZ W ->
        (2560, 2, 2560) with the consolidate_indices layout (contract_mps_mpo).
    Prints one JSON line with each kernel's average time, algorithmic work and
    rate from the library's launch timers."""
    import numpy as np
    import torch
    import tncontract_amd as tn
    from tncontract_amd import _lib as L
    from tncontract_amd import tensor as tsr
    from tncontract_amd.onedim import xxz_mpo
    torch.cuda.set_device(0)
    rng = np.random.default_rng(args.seed)
    chi, D, d = args.chi, 5, 2
    cd = chi * D
    R64 = torch.from_numpy(rng.standard_normal((cd, cd))).cuda()
    S64 = torch.from_numpy(rng.standard_normal((cd, d, cd))).cuda()
    B = tn.Tensor(rng.standard_normal((d, chi, chi)), ["phys", "left", "right"])
    W = xxz_mpo(3)[1]
    reps = max(args.steps, 3)
    out = {}

    def timed(tag, fn, flops, nbytes):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / reps
        out[tag] = {"ms": dt * 1e3, "flops": flops, "bytes": nbytes,
                    "tflops": flops / dt / 1e12 if flops else None,
                    "gbs": nbytes / dt / 1e9}
        log("%s: %.3f ms" % (tag, dt * 1e3))

    fl = 2.0 * cd * cd * d * cd
    for prec in ("native", "fp32", "bf16x3"):
        tag = {"native": "gemm_f64", "fp32": "gemm_f32", "bf16x3": "gemm_bf16x3"}[prec]
        timed(tag, lambda: tsr._contract_data(R64, [1], S64, [0], prec), fl,
              8.0 * (cd * cd + 2 * cd * d * cd))
    Sv = S64.permute(1, 0, 2)
    timed("permute", lambda: L.materialize(Sv), 0.0, 16.0 * S64.numel())
    timed("direct", lambda: tn.contract(B, W, "phys", "physin"), 2.0 * d * cd * d * cd,
          8.0 * (B.data.numel() + cd * d * cd))
    rec = {"metric": "hot-kernel rates at the chi=%d sweep's shapes" % chi, "value": None,
           "unit": None, "n_gpus": 1, "steps": reps, "warmup": 1, "higher_is_better": True,
           "data": "synthetic (seeded N(0,1))",
           "config": {"workload": "contraction GEMMs (%d x %d) . (%d x %d x %d), permute of a "
                                  "(%d, %d, %d) site, MPO application of a (%d, %d, %d) site"
                                  % (cd, cd, cd, d, cd, cd, d, cd, d, chi, chi)},
           "kernels": out}
    return rec


def config5_workload(args):
    """Config 5: XXZ MPO (D=5) applied to a seeded random MPS N=128, d=2,
    chi=1024 (fused bond 5120) and compressed back to chi=1024, four times in
    a row (each sweep's output is the next one's input), bf16x3 contractions
    with FP64 eigensolver rotations and isometries.  One step = one sweep.
    Reports the peak device memory and, for the first sweep, the state
    distance to the FP32-contraction sweep.  The oracle is not run (about
    209 TFLOP per sweep on the host)."""
    import torch
    import tncontract_amd as tn
    from tncontract_amd import _lib as L
    from tncontract_amd.onedim import random_mps_arrays, xxz_mpo
    from tncontract_amd.onedim.structured import apply_mpo_compress, SweepStats
    torch.cuda.set_device(0)
    N, d, chi = 128, 2, 1024
    arrs = random_mps_arrays(N, d, chi, seed=args.seed)
    psi = tn.onedim.MatrixProductState([tn.Tensor(a, ["phys", "left", "right"]) for a in arrs])
    mpo = xxz_mpo(N)
    torch.cuda.synchronize()
    torch.cuda.reset_peak_memory_stats()
    log("config5: first sweep (bf16x3), then the FP32 one for the distance")
    t0 = time.perf_counter()
    first = apply_mpo_compress(psi, mpo, chi=chi, threshold=1e-15, precision="bf16x3")
    torch.cuda.synchronize()
    t_first = time.perf_counter() - t0
    ref = apply_mpo_compress(psi, mpo, chi=chi, threshold=1e-15, precision="fp32")
    ip = tn.onedim.inner_product_mps
    nf = float(ip(ref, ref))
    dist = (float(ip(first, first)) + nf - 2 * float(ip(first, ref))) / nf
    del ref
    steps = max(args.steps, 1)
    log("config5: %d further sweeps" % steps)
    cur = first
    st = SweepStats()
    t0 = time.perf_counter()
    for k in range(steps):
        if k == steps - 1:
            L.prof_collect("gemm_f64", reset=True)
            L.prof_enable(True)
        cur = apply_mpo_compress(cur, mpo, chi=chi, threshold=1e-15, precision="bf16x3",
                                 stats=st if k == steps - 1 else None)
        torch.cuda.synchronize()
        log("config5: sweep %d done, %.1f s" % (k + 2, time.perf_counter() - t0))
    L.prof_enable(False)
    prof = {tag: L.prof_collect(tag) for tag in L.PROF_TAGS}
    L.prof_collect("gemm_f64", reset=True)
    dt = (time.perf_counter() - t0) / steps
    peak = torch.cuda.max_memory_allocated()
    roof = roofline_from_prof(prof, args.traffic)
    rec = {"metric": "MPO.MPS compress sweeps/s at chi=1024 (config 5)", "value": 1.0 / dt,
           "unit": "sweeps/s", "n_gpus": 1, "steps": steps, "warmup": 1,
           "ms_per_step": dt * 1e3, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "bf16x3+f64",
           "data": "synthetic (seeded random MPS, i.i.d. N(0,1), normalised; XXZ MPO)",
           "config": {"workload": "XXZ MPO D=5 x random MPS N=128 d=2 chi=1024 -> compress "
                                  "chi=1024, %d sweeps after the first, each on the previous "
                                  "output; bf16x3 contractions" % steps},
           "first_sweep_s": t_first, "peak_device_memory_gb": peak / 1e9,
           "first_sweep_state_dist_bf16x3_vs_fp32": dist,
           "eig_jacobi_sweeps_mean": sum(st.sweeps.values()) / max(len(st.sweeps), 1),
           "structured_fallback_bond": st.fallback,
           "kernel_ms_profiled_step": {t: round(v[0], 3) for t, v in prof.items() if v[1]},
           "roofline": roof,
           "cpu_baseline": None,
           "cpu_baseline_note": "not run: about 209 TFLOP of FP64 per sweep on the host"}
    return rec


def learned_workload(args):
    """Config 4: the truncating sweep of svd_compress on a d=4 chain whose
    input bond is 1024 (bond matrices up to 2048 x 1024), compressed to
    chi=512 at sigma/sigma_0 > 1e-8, FP32 data in FP64 arithmetic.  The rank
    model is trained on the GPU on the spectra of full-SVD sweeps of other
    states of the family and of config 1; one step = one learned-rank sweep;
    the same sweep with the full Jacobi SVD at every bond (the method without
    the network) is timed beside it."""
    import numpy as np
    import torch
    from tncontract_amd import _lib as L
    from tncontract_amd import learned as lr
    torch.cuda.set_device(0)
    chi, thr = 512, 1e-8
    log("learned: training sweeps")
    t0 = time.perf_counter()
    samples = []
    for seed in (1, 2, 3, 4):
        prev, cur, _ = lr.decaying_pair(seed)
        samples.append(lr.sweep_samples(prev, cur, chi, thr)[:2])
    for seed in (1, 2):
        prev, cur = lr.flat_pair(seed, 20, 2, 32)
        samples.append(lr.sweep_samples(prev, cur, 16, 1e-10)[:2])
    t_data = time.perf_counter() - t0
    t0 = time.perf_counter()
    model, losses = lr.train_rank_model(samples)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    prev, cur, taus = lr.decaying_pair(99 + args.seed)
    X, T, spectra_prev, needed_cur = lr.sweep_samples(prev, cur, chi, thr)
    pred = model.predict_rank(X, chi)
    rank_err = float(np.mean(np.abs(pred - T * chi)) / chi)
    state = lr._state(cur)
    torch.cuda.synchronize()

    def learned_step(stats=None):
        return lr.learned_svd_compress(state, chi, thr, model, spectra_prev, stats=stats)

    def full_step():
        return lr.record_sweep(state, chi, thr)[0]

    log("learned warmup")
    for _ in range(args.warmup):
        learned_step()
        full_step()
    torch.cuda.synchronize()
    log("learned timed x%d" % args.steps)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out = learned_step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ref = full_step()
    torch.cuda.synchronize()
    dt_full = (time.perf_counter() - t0) / args.steps
    st = lr.LearnedSweepStats()
    L.prof_collect("gemm_f64", reset=True)
    L.prof_enable(True)
    learned_step(st)
    torch.cuda.synchronize()
    L.prof_enable(False)
    prof = {tag: L.prof_collect(tag) for tag in L.PROF_TAGS}
    L.prof_collect("gemm_f64", reset=True)
    roof = roofline_from_prof(prof, args.traffic)
    ip = __import__("tncontract_amd").onedim.inner_product_mps
    nn = float(ip(ref, ref))
    err = float((float(ip(out, out)) + nn - 2 * float(ip(out, ref))) / nn)
    cpu, err_oracle = None, None
    if not args.no_cpu:
        from oracle import tn_oracle as o
        log("cpu baseline: oracle svd_compress of the same state")
        r = o.OMPS([o.OTensor(a.astype(np.float64), ["phys", "left", "right"]) for a in cur])
        t1 = time.perf_counter()
        r.svd_compress(chi=chi, threshold=thr)
        tc = time.perf_counter() - t1
        g = o.OMPS([o.OTensor(t.numpy(), t.labels) for t in out], out.left_label,
                   out.right_label, out.phys_label)
        nr = o.inner_product_mps(r, r)
        err_oracle = float((o.inner_product_mps(g, g) + nr - 2 * o.inner_product_mps(g, r)) / nr)
        cpu = {"value": 1.0 / tc, "unit": "sweeps/s", "cores": len(os.sched_getaffinity(0)),
               "kind": "port", "sample": "NumPy FP64 oracle svd_compress of the same state once "
                                         "(%.2f s)" % tc}
    rec = {"metric": "learned-rank compression sweeps/s (config 4)", "value": 1.0 / dt,
           "unit": "sweeps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": dt * 1e3, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64 (FP32 data)",
           "data": "synthetic (seeded d=4 MPS, input bond 1024, bond spectra ~exp(-k/tau), "
                   "tau ~ U[2,16])",
           "config": {"workload": "svd_compress to chi=512 at sigma/sigma_0 > 1e-8, N=16, d=4, "
                                  "input bond 1024 (bond matrices up to 2048 x 1024); rank model "
                                  "trained on %d sweep samples" % sum(len(x[1]) for x in samples)},
           "full_svd_sweeps_per_s": 1.0 / dt_full, "speedup_vs_full_svd_sweep": dt_full / dt,
           "state_err_vs_full_svd_sweep": err, "state_err_vs_oracle": err_oracle,
           "bond_dims_equal_full_svd": out.bonddims() == ref.bonddims(),
           "fallbacks": st.fallbacks, "paths": {str(b): p for b, p in sorted(st.path.items())},
           "predictor_rank_error_mean_over_chi": rank_err,
           "columns_factored": st.work_cols, "columns_full_svd": st.full_cols,
           "train": {"samples": int(sum(len(x[1]) for x in samples)), "data_s": t_data,
                     "train_s": t_train, "loss_first": losses[0], "loss_last": losses[-1]},
           "kernel_ms_profiled_step": {t: round(v[0], 3) for t, v in prof.items() if v[1]},
           "roofline": roof, "cpu_baseline": cpu}
    return rec


def xxz_bond_gate(tau, Jz=1.0, Jxy=1.0):
    """exp(-tau h) of the XXZ bond Hamiltonian h = Jz Sz Sz + Jxy/2 (S+S- +
    S-S+) (spin 1/2), as a (o0, o1, i0, i1) array."""
    import numpy as np
    sz = np.diag([0.5, -0.5])
    sp = np.array([[0.0, 1.0], [0.0, 0.0]])
    h = Jz * np.kron(sz, sz) + 0.5 * Jxy * (np.kron(sp, sp.T) + np.kron(sp.T, sp))
    w, v = np.linalg.eigh(h)
    return ((v * np.exp(-tau * w)) @ v.T).reshape(2, 2, 2, 2)


def tebd_workload(args):
    """Two-site gates/s of MatrixProductStateCanonical.apply_gate
    (onedim_core.py:1100-1186): one step = one imaginary-time XXZ Trotter
    sweep (every even bond, then every odd bond) of a canonical chain,
    truncating each updated bond to chi."""
    import numpy as np
    import torch
    import tncontract_amd as tn
    from tncontract_amd import _lib as L
    from tncontract_amd.onedim import random_mps_arrays
    torch.cuda.set_device(0)
    N, d, chi = args.tebd_N, 2, args.tebd_chi
    arrs = random_mps_arrays(N, d, chi, seed=args.seed)
    G = xxz_bond_gate(0.05)
    gate = tn.Tensor(G, ["o0", "o1", "i0", "i1"])
    base = tn.onedim.canonical_form_mps(
        tn.onedim.MatrixProductState([tn.Tensor(a, ["phys", "left", "right"]) for a in arrs]),
        threshold=1e-14, normalise=True)
    even, odd = list(range(0, N - 1, 2)), list(range(1, N - 1, 2))
    bonds = even + odd
    torch.cuda.synchronize()

    def step(c):
        # the two layers of the sweep, each bond's SVD batched with its layer
        c.apply_gate_layer(gate, even, chi=chi, threshold=1e-14)
        c.apply_gate_layer(gate, odd, chi=chi, threshold=1e-14)
        return c

    def step_sequential(c):
        for k in bonds:
            c.apply_gate(gate, k, chi=chi, threshold=1e-14)
        return c

    log("tebd: N=%d chi=%d, %d gates per sweep" % (N, chi, len(bonds)))
    for _ in range(args.warmup):
        step(base.copy())
        step_sequential(base.copy())
    torch.cuda.synchronize()
    runs = [base.copy() for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for c in runs:
        step(c)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    seq = base.copy()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    step_sequential(seq)
    torch.cuda.synchronize()
    dt_seq = time.perf_counter() - t0
    mid = seq.singular_site(N // 2)
    lay = np.diag(runs[0][mid].numpy())
    sq = np.diag(seq[mid].numpy())
    layer_vs_seq = float(np.max(np.abs(lay - sq)) / sq[0]) if len(lay) == len(sq) else float("nan")
    L.prof_collect("gemm_f64", reset=True)
    L.prof_enable(True)
    step(base.copy())
    torch.cuda.synchronize()
    L.prof_enable(False)
    prof = {tag: L.prof_collect(tag) for tag in L.PROF_TAGS}
    L.prof_collect("gemm_f64", reset=True)
    roof = roofline_from_prof(prof, args.traffic)
    if roof["kernel"].startswith("jac"):
        roof["note"] = ("batched Jacobi work is counted on the device: rows and eigensolves of "
                        "the bonds and block pairs still rotating")
    cpu, lam_err = None, None
    if not args.no_cpu:
        from oracle import tn_oracle as o
        log("cpu baseline: oracle on the same sweep")
        oc = o.OMPSC([o.OTensor(t.numpy(), list(t.labels)) for t in base], base.left_label,
                     base.right_label, base.phys_label)
        og = o.OTensor(G, ["o0", "o1", "i0", "i1"])
        t1 = time.perf_counter()
        for k in bonds:
            oc.apply_gate(og, k, chi=chi, threshold=1e-14)
        tc = time.perf_counter() - t1
        mid = runs[0].singular_site(N // 2)
        lg, lr = np.diag(runs[0][mid].numpy()), np.diag(oc[mid].data)
        lam_err = float(np.max(np.abs(lg - lr)) / lr[0]) if len(lg) == len(lr) else float("nan")
        cpu = {"value": len(bonds) / tc, "unit": "gates/s", "cores": len(os.sched_getaffinity(0)),
               "kind": "port", "sample": "NumPy FP64 oracle, the same Trotter sweep once "
                                         "(%.2f s)" % tc}
    rec = {"metric": "canonical-MPS two-site gates/s (apply_gate_layer)", "value":
           len(bonds) / dt, "unit": "gates/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": dt * 1e3, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic (seeded random MPS in canonical form, normalised)",
           "config": {"workload": "imaginary-time XXZ Trotter sweep (tau=0.05, even then odd "
                                  "bonds), canonical MPS N=%d d=2 chi=%d, truncation chi=%d; "
                                  "value: apply_gate_layer (one batched FP64 SVD launch set per "
                                  "layer)" % (N, chi, chi)},
           "sequential_apply_gate_gates_per_s": len(bonds) / dt_seq,
           "layer_vs_sequential_mid_lambda_rel_diff": layer_vs_seq,
           "mid_bond_lambda_rel_err_vs_oracle": lam_err,
           "kernel_ms_profiled_step": {t: round(v[0], 3) for t, v in prof.items() if v[1]},
           "roofline": roof, "cpu_baseline": cpu}
    return rec


def main():
    args = parse()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print("bench.py: measures one GPU (the sweep does not shard within a chain); "
              "--gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)
    others = {"batched": batched_workload, "variational": variational_workload,
              "peps": peps_workload, "tebd": tebd_workload, "learned": learned_workload,
              "config5": config5_workload, "kernels": kernels_workload}
    if args.workload in others:
        print(json.dumps(others[args.workload](args)), flush=True)
        return
    print(json.dumps(chain_workload(args)), flush=True)


def chain_workload(args):
    """The headline: MPO.MPS compress sweeps/s at chi=512 (see the module
    docstring); the batched (config 3) and learned-rank (config 4) workloads
    run after it with their own steps and are reported inside its record."""
    import numpy as np
    import torch

    import tncontract_amd as tn
    from tncontract_amd import _lib as L
    from tncontract_amd.onedim import random_mps_arrays, xxz_mpo
    from tncontract_amd.onedim.structured import apply_mpo_compress, SweepStats

    torch.cuda.set_device(0)
    N, d, chi = args.N, 2, args.chi
    arrs = random_mps_arrays(N, d, chi, seed=args.seed)
    psi = tn.onedim.MatrixProductState([tn.Tensor(a, ["phys", "left", "right"]) for a in arrs])
    mpo = xxz_mpo(N)
    torch.cuda.synchronize()

    def fast():
        return apply_mpo_compress(psi, mpo, chi=chi, threshold=1e-15, precision=args.precision,
                                  eig_tol=args.eig_tol)

    def plain():
        out = tn.onedim.contract_mps_mpo(psi, mpo)
        out.svd_compress(chi=chi, threshold=1e-15)
        return out

    # ---- fast path (the metric)
    log("fast path warmup")
    for _ in range(args.warmup):
        fast()
    torch.cuda.synchronize()
    log("fast path timed x%d" % args.steps)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        res_fast = fast()
    torch.cuda.synchronize()
    dt_fast = (time.perf_counter() - t0) / args.steps
    log("fast: %.3f s/sweep" % dt_fast)

    # ---- one profiled fast sweep: per-kernel launch timing (HIP events)
    L.prof_collect("gemm_f64", reset=True)
    L.prof_enable(True)
    st = SweepStats()
    apply_mpo_compress(psi, mpo, chi=chi, threshold=1e-15, precision=args.precision, stats=st,
                       eig_tol=args.eig_tol)
    torch.cuda.synchronize()
    L.prof_enable(False)
    prof = {tag: L.prof_collect(tag) for tag in L.PROF_TAGS}
    L.prof_collect("gemm_f64", reset=True)
    roof = roofline_from_prof(prof, args.traffic, MEASURED_TRAFFIC)
    kernel_ms = {t: round(v[0], 3) for t, v in prof.items() if v[1]}
    sweep_roof = sweep_roofline(prof, dt_fast)

    # ---- plain FP64 path (the denominator) and the error between the two
    dt_plain, err = None, None
    if not args.no_plain:
        log("plain path warmup")
        plain()
        torch.cuda.synchronize()
        log("plain path timed x%d" % args.plain_steps)
        t0 = time.perf_counter()
        for _ in range(args.plain_steps):
            res_plain = plain()
        torch.cuda.synchronize()
        dt_plain = (time.perf_counter() - t0) / args.plain_steps
        log("plain: %.3f s/sweep" % dt_plain)
        ip = tn.onedim.inner_product_mps
        nn = float(ip(res_plain, res_plain))
        err = float((float(ip(res_fast, res_fast)) + nn - 2 * float(ip(res_fast, res_plain))) / nn)
        if res_fast.bonddims() != res_plain.bonddims():
            err = float("nan")

    cpu, check = None, None
    if not args.no_cpu:
        log("cpu baseline: the oracle's full sweep of a 22-site chain, and the GPU paths "
            "checked against it")
        cpu, check = cpu_baseline_sweep(args)

    rec = {
        "metric": METRIC,
        "value": 1.0 / dt_fast,
        "unit": "sweeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt_fast * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "f64", "fp32": "f32+f64", "bf16x3": "bf16x3+f64"}[args.precision],
        "data": "synthetic (seeded random MPS, i.i.d. N(0,1), normalised; XXZ MPO)",
        "config": {"workload": "XXZ MPO D=5 x random MPS N=%d d=%d chi=%d -> compress chi=%d"
                               % (N, d, chi, chi),
                   "fast_path": ("structured apply+compress (no chi*D chain, no QR sweep), %s "
                                 "contractions, FP64 Cholesky + one-sided Jacobi eigensolver of "
                                 "each bond's Gram matrix, FP64 isometries; the learned-rank "
                                 "randomized SVD is not in this timed path (every bond of this "
                                 "flat-spectrum workload keeps half of its Gram's rank, where a "
                                 "sketch is no smaller than the full solve) and is measured by "
                                 "--workload learned" % args.precision)},
        "plain_fp64_sweeps_per_s": (1.0 / dt_plain) if dt_plain else None,
        "speedup_vs_plain_fp64": (dt_plain / dt_fast) if dt_plain else None,
        "state_err_vs_plain_fp64": err,
        "kernel_ms_profiled_sweep": kernel_ms,
        "eig_jacobi_sweeps_mean": (sum(st.sweeps.values()) / max(len(st.sweeps), 1)),
        "structured_fallback_bond": st.fallback,
        "sweep_roofline": sweep_roof,
        "roofline": roof,
        "cpu_baseline": cpu,
        "oracle_check": check,
    }
    if not args.no_extra:
        ex = argparse.Namespace(**vars(args))
        ex.steps, ex.warmup, ex.traffic, ex.seed = 2, 1, None, 0
        keys = ("value", "unit", "ms_per_step", "config", "roofline", "cpu_baseline")
        log("config 3: batched truncation")
        b = batched_workload(ex)
        rec["batched_config3"] = dict({k: b[k] for k in keys}, kept_vs_lapack_sample=b.get(
            "kept_equal_lapack_sample"), sv_rel_err_sample=b.get("sv_rel_err_sample"),
            jacobi_sweeps_max=b.get("jacobi_sweeps_max"))
        log("config 4: learned-rank sweep")
        lw = learned_workload(ex)
        rec["learned_config4"] = dict({k: lw[k] for k in keys}, **{
            k: lw[k] for k in ("full_svd_sweeps_per_s", "speedup_vs_full_svd_sweep",
                               "state_err_vs_oracle", "fallbacks", "paths",
                               "predictor_rank_error_mean_over_chi", "columns_factored",
                               "columns_full_svd")})
    return rec


if __name__ == "__main__":
    main()
