"""Train the learned preconditioner correction on the GPU and evaluate it.

    python -m earthbox_gpu.train_learned [--out PATH] [--steps N] [--fields K]

Training systems: layered random log-viscosity fields on 1001 x 501 nodes
(1500 x 600 km), seeds 40, 42, 43, ...; samples (default) from the first
Krylov vectors of each system's flexible GMRES iteration with their exact
errors, or from random smooth error fields, through the plain FP32 V-cycle
(ebg_cnn_make_sample), cropped to 128 x 64 cells; AdamW, batch 32.  Evaluation: FGMRES to 1e-8 on the held-out
seed 41 and on seed 2 (the benchmark system), with and without the
correction, reporting iterations, time and accepted corrections.
"""
from __future__ import annotations

import argparse
import json
import sys
import time


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=None)
    ap.add_argument("--steps", type=int, default=1500)
    ap.add_argument("--fields", type=int, default=8)
    ap.add_argument("--samples", type=int, default=24)
    ap.add_argument("--crops", type=int, default=16)
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--xnum", type=int, default=1001)
    ap.add_argument("--ynum", type=int, default=501)
    ap.add_argument("--eval-seeds", default="41,2")
    ap.add_argument("--data", choices=("krylov", "smooth"), default="krylov",
                    help="krylov: the preconditioner's Krylov inputs with exact errors; "
                         "smooth: random smooth error fields")
    args = ap.parse_args(argv)

    import torch
    import earthbox_gpu
    from earthbox_gpu import learned, workloads
    from earthbox_gpu.stokes import StokesSolverGPU, bc_mirror_factors

    def log(msg):
        print(msg, flush=True)

    ctx = earthbox_gpu.GPUContext(0)
    net = learned.LearnedCorrection(ctx, seed=1)
    log(f"parameters: {net.num_params} stored, {learned.effective_param_count()} effective")
    seeds = [40] + list(range(42, 42 + args.fields - 1))
    t0 = time.perf_counter()
    if args.data == "krylov":
        X, Y = learned.build_krylov_dataset(ctx, net, seeds, args.samples, args.crops, xnum=args.xnum,
                                            ynum=args.ynum, log=log)
    else:
        X, Y = learned.build_dataset(ctx, net, seeds, args.samples, args.crops, xnum=args.xnum,
                                     ynum=args.ynum, log=log)
    log(f"dataset: {X.shape[0]} crops in {time.perf_counter() - t0:.1f} s")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    hist = learned.train(net, X, Y, args.steps, batch=args.batch, lr=args.lr, log=log)
    torch.cuda.synchronize()
    log(f"training: {args.steps} steps in {time.perf_counter() - t0:.1f} s, "
        f"loss {hist[0]:.4f} -> {sum(hist[-50:]) / min(50, len(hist)):.4f}")
    if args.out:
        net.save(args.out)
        log(f"saved {args.out}")
    del X, Y
    results = {}
    sol = StokesSolverGPU(ctx)
    for s in [int(v) for v in args.eval_seeds.split(",") if v]:
        prob = workloads.layered_random(args.xnum, args.ynum, 1.5e6 * (args.xnum - 1) / 1000,
                                        6.0e5 * (args.ynum - 1) / 500, seed=s)
        sol.set_system(prob)
        bcd = bc_mirror_factors(prob)
        row = {}
        for name, kw in (("mixed", dict(precision=1)), ("learned1_mixed", dict(precision=1, learned=1)),
                         ("learned2_mixed", dict(precision=1, learned=2)),
                         ("learned3_mixed", dict(precision=1, learned=3)),
                         ("plain_fp64", dict(precision=0))):
            sol.solve_system(rtol=1e-8, max_iterations=2, bc_d=bcd, **kw)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            _, st = sol.solve_system(rtol=1e-8, max_iterations=1000, bc_d=bcd, **kw)
            torch.cuda.synchronize()
            row[name] = {"seconds": time.perf_counter() - t1, "iterations": st["iterations"],
                         "relres": st["relres"], "converged": st["converged"],
                         "learned_accepted": st["learned_accepted"]}
            log(f"seed {s} {name}: {row[name]}")
        results[str(s)] = row
    print(json.dumps({"eval": results}), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
