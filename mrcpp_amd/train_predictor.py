"""Train the refinement predictor on refine labels of the plain adaptive apply.

    python -m mrcpp_amd.train_predictor [--train 200] [--held 50] [--epochs 30] [--out mrcpp_amd/share/predictor_k7.npy]

Trees: mrcpp_amd.workloads.predictor_sources (order 7, prec 1e-6, 4..32
Gaussians, Poisson and Helmholtz).  Training samples come from batched plain
applies (every output node beyond the first step: 12 features, refined or
not); the network is trained on the device by the library's own gradient and
Adam kernels.  The held-out trees are then applied plainly and with the
predictor, batched, and the report (JSON on stdout) gives device passes,
node-set identity, error against the plain result and wall time of both.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def build_trees(mr, mra, specs):
    trees = []
    ops = {}
    for sp in specs:
        f = mr.FunctionTree(mra)
        g = mr.GaussExp.from_sources(sp["sources"])
        mr.build_grid(f, g)
        mr.project(sp["prec"], f, g)
        key = (sp["mu"], sp["prec"])
        if key not in ops:
            ops[key] = mr.PoissonOperator(mra, sp["prec"]) if sp["mu"] is None else mr.HelmholtzOperator(mra, sp["mu"], sp["prec"])
        trees.append((sp["prec"], ops[key], f, np.asarray(sp["sources"].centers)))
    return trees


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=200)
    ap.add_argument("--held", type=int, default=50)
    ap.add_argument("--chunk", type=int, default=16)
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--threshold", type=float, default=0.5)
    ap.add_argument("--out", default=os.path.join(HERE, "share", "predictor_k7.npy"))
    ap.add_argument("--evaluate", default=None, help="skip training: evaluate these parameters on the held-out trees")
    args = ap.parse_args(argv)
    sys.path.insert(0, os.path.dirname(HERE))
    import mrcpp_amd as mr
    from mrcpp_amd import workloads as W

    specs = W.predictor_sources(n_trees=args.train + args.held)
    mra = mr.MultiResolutionAnalysis((-32, 32), 7)
    t0 = time.time()
    xs, ys = [], []
    for c in range(0, 0 if args.evaluate else args.train, args.chunk):
        trees = build_trees(mr, mra, specs[c:min(args.train, c + args.chunk)])
        outs = [mr.FunctionTree(mra) for _ in trees]
        _, x, y = mr.apply_predicted([(t[0], o, t[1], t[2]) for t, o in zip(trees, outs)], centers=[t[3] for t in trees], collect=True)
        xs.append(x)
        ys.append(y)
        print("training trees %d..%d: %d samples" % (c, c + len(trees) - 1, len(y)), file=sys.stderr, flush=True)
        del trees, outs
    pred = mr.RefinePredictor()
    if args.evaluate:
        pred.params = np.load(args.evaluate)
        rep = dict(params=args.evaluate)
    else:
        X, Y = np.concatenate(xs), np.concatenate(ys)
        t_data = time.time() - t0
        t0 = time.time()
        losses = pred.train(X, Y, epochs=args.epochs, batch=4096, lr=2e-3, pos_weight=3.0, seed=7)
        t_train = time.time() - t0
        os.makedirs(os.path.dirname(args.out), exist_ok=True)
        np.save(args.out, pred.params)
        rep = dict(samples=int(len(Y)), refined_fraction=float(Y.mean()), losses=[float(v) for v in losses], seconds_data=t_data,
                   seconds_train=t_train)
    pred.threshold = args.threshold
    rep["threshold"] = args.threshold
    # held-out trees, batched: plain against predicted
    held = build_trees(mr, mra, specs[args.train:args.train + args.held])
    # warm-up: operator tables and input mirrors are built on first use
    warm = [mr.FunctionTree(mra) for _ in held]
    mr.apply_predicted([(t[0], o, t[1], t[2]) for t, o in zip(held, warm)], predictor=pred, centers=[t[3] for t in held])
    del warm
    res = {}
    for name, p in (("plain", None), ("predicted", pred)):
        mr.device_synchronize()
        t0 = time.time()
        outs = [mr.FunctionTree(mra) for _ in held]
        st = mr.apply_predicted([(t[0], o, t[1], t[2]) for t, o in zip(held, outs)], predictor=p, centers=[t[3] for t in held])
        mr.device_synchronize()
        res[name] = dict(seconds=time.time() - t0, outs=outs, st=st)
    same = 0
    num = den = 0.0
    for a, b in zip(res["plain"]["outs"], res["predicted"]["outs"]):
        ai, ac, _ = a.export()
        bi, bc, _ = b.export()
        if np.array_equal(ai, bi):
            same += 1
            num += float(np.sum((bc - ac) ** 2))
            den += float(np.sum(ac ** 2))
    sp, sq = res["plain"]["st"], res["predicted"]["st"]
    rep["held_out"] = dict(
        trees=len(held), same_node_set=same, rel_l2_vs_plain=(num / den) ** 0.5 if den else None,
        passes_plain=int(max(s["n_levels"] for s in sp)), passes_predicted=int(max(s["n_levels"] for s in sq)),
        seconds_plain=res["plain"]["seconds"], seconds_predicted=res["predicted"]["seconds"],
        predicted_nodes=int(sum(s["n_predicted"] for s in sq)), cropped_nodes=int(sum(s["n_cropped"] for s in sq)),
        fallback_nodes=int(sum(s["n_fallback"] for s in sq)), output_nodes=int(sum(s["n_output_nodes"] for s in sp)))
    print(json.dumps(rep))


if __name__ == "__main__":
    main()
