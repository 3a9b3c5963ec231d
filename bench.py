"""MPO.MPS compress sweeps/s on one MI355X.

Workload (the configuration the headline metric is quoted on): XXZ MPO
(D=5, Jz=1, Jxy=1, hz=0.5) applied to a seeded random MPS (N=64, d=2,
bonds [1,512,...,512,1], i.i.d. N(0,1) FP64 entries, normalised) and
compressed back to chi=512 (threshold 1e-15).  One step = one sweep =
contract_mps_mpo + svd_compress on the same resident input state.

Paths:
  plain -- the reference algorithm in FP64 on the GPU (contract_mps_mpo,
           QR left sweep, truncating SVD right sweep).
  fast  -- (when available) the mixed-precision / batched / learned-rank sweep.

Prints ONE JSON line (rank 0, N=1).
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


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=2)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--path", default="plain", choices=["plain"])
    p.add_argument("--N", type=int, default=64)
    p.add_argument("--chi", type=int, default=512)
    p.add_argument("--seed", type=int, default=0)
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline sample")
    return p.parse_args()


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (every box here has one GPU); "
              "--gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    import tncontract_amd as tn
    from tncontract_amd.onedim import random_mps_arrays, xxz_mpo

    torch.cuda.set_device(0)
    N, d, chi = args.N, 2, args.chi
    arrs = random_mps_arrays(N, d, chi, seed=args.seed)
    psi = tn.onedim.MatrixProductState([tn.Tensor(a, ["phys", "left", "right"]) for a in arrs])
    mpo = xxz_mpo(N)
    torch.cuda.synchronize()

    def sweep():
        out = tn.onedim.contract_mps_mpo(psi, mpo)
        out.svd_compress(chi=chi, threshold=1e-15)
        return out

    for _ in range(args.warmup):
        sweep()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        res = sweep()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    rec = {
        "metric": METRIC,
        "value": 1.0 / dt,
        "unit": "sweeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded random MPS, i.i.d. N(0,1), normalised)",
        "config": {"workload": "XXZ MPO (D=5) x random MPS N=%d d=%d chi=%d -> compress chi=%d"
                               % (N, d, chi, chi), "path": args.path,
                   "bonddims_out": max(res.bonddims())},
    }
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
