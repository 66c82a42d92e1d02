"""Trains the sign-risk proposer on the GPU with the library's own kernels.

Data: seeded random vessels (perturbed cups / cup+torus / torus pairs,
50k-300k triangles, log-uniform), each framed by the reference rule at a
voxel size giving 100-300 voxels across.  For every candidate voxel of the
learned method (band distance within the jitter reach) the library returns
its 12 features and scanline sign; the label is 1 when the exact winding
kernel (compensated FP32, FP64 near 1/2) disagrees with the scanline sign.

Training: weighted binary cross-entropy, Adam, minibatches, all through
``rhz_mlp_grad`` / ``rhz_mlp_adam`` (BF16 matrix-core forward and backward).
The result is written as ``csrc/proposer_weights.inc`` (the built-in
parameter blob the library compiles in) plus a JSON summary; a named variant
(``--variant noflag --drop-feature 10``: a network that never sees the
certificate flag) is written as ``proposer_<variant>.npy`` instead.

Usage (needs a GPU):
    python rhizome_v24.15_amd/proposer_train.py --meshes 200 --out DIR
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
N_FEATURES = 12
N_H = 64


def _load():
    sys.path.insert(0, os.path.dirname(HERE))
    import __graft_entry__ as entry  # noqa: E402

    return entry.load_package(), entry.load_geometry()


def init_blob(n_blob, n_param, seed, mean, istd):
    """He-uniform weights, zero biases; normalisation statistics appended."""
    rng = np.random.default_rng(seed)
    b = np.zeros(n_blob, dtype=np.float32)
    o = 0
    for fan_in, fan_out in ((N_FEATURES, N_H), (N_H, N_H), (N_H, N_H), (N_H, 1)):
        lim = math.sqrt(6.0 / fan_in)
        b[o:o + fan_in * fan_out] = rng.uniform(-lim, lim, fan_in * fan_out)
        o += fan_in * fan_out + fan_out  # weights, then zero biases
    assert o == n_param
    b[n_param:n_param + N_FEATURES] = mean
    b[n_param + N_FEATURES:n_param + 2 * N_FEATURES] = istd
    return b


def mesh_samples(rhz, geom, torch, seed):
    """Features and labels of one seeded random vessel (GPU tensors)."""
    rng = np.random.default_rng(seed)
    target = int(math.exp(rng.uniform(math.log(50_000), math.log(300_000))))
    v, f = geom.random_vessel(seed, target, tri_bounds=(50_000, 300_000))
    span = float(np.max(v.max(0) - v.min(0)))
    h = float(np.float32(span / rng.uniform(100.0, 300.0)))
    mn, mx = geom.frame_grid(v, h)
    tris = rhz.build_triangles(v, f)
    dims = [mx[i] - mn[i] + 1 for i in range(3)]
    V = int(np.prod(np.array(dims, dtype=np.int64)))
    d_t = torch.from_numpy(tris).cuda()
    n = rhz.Proposer.candidates(d_t, len(tris), h, mn, mx, None, None, 0)
    d_v = torch.empty(n, dtype=torch.int64, device="cuda")
    d_f = torch.empty((n, N_FEATURES), dtype=torch.float32, device="cuda")
    d_s = torch.empty(V, dtype=torch.uint8, device="cuda")
    rhz.Proposer.candidates(d_t, len(tris), h, mn, mx, d_v, d_f, n, d_scan=d_s)
    d_w = torch.empty(n, dtype=torch.float64, device="cuda")
    rhz.Stages.winding_voxels(d_t, len(tris), h, mn, dims, d_v, n, d_w)
    exact = (d_w.abs() >= 0.5).to(torch.uint8)
    y = (exact != d_s[d_v]).to(torch.float32)
    return d_f, y, {"seed": seed, "tris": len(tris), "h": h, "voxels": V, "candidates": n,
                    "positives": int(y.sum().item())}


def generate(rhz, geom, torch, seeds, log):
    xs, ys, meta = [], [], []
    for k, s in enumerate(seeds):
        x, y, m = mesh_samples(rhz, geom, torch, s)
        xs.append(x)
        ys.append(y)
        meta.append(m)
        if k % 20 == 0:
            log(f"mesh {k}/{len(seeds)}: {m}")
    return torch.cat(xs), torch.cat(ys), meta


def evaluate(rhz, torch, d_blob, X, Y):
    z = torch.empty(len(X), dtype=torch.float32, device="cuda")
    rhz.Proposer.forward(d_blob, X, len(X), z)
    pred = z >= 0
    pos = Y > 0.5
    tp = int((pred & pos).sum().item())
    fp = int((pred & ~pos).sum().item())
    fn = int((~pred & pos).sum().item())
    return {"samples": len(X), "positives": int(pos.sum().item()), "predicted": tp + fp, "true_pos": tp,
            "false_neg": fn, "precision": tp / max(tp + fp, 1), "recall": tp / max(tp + fn, 1)}


def train(rhz, torch, X, Y, Xv, Yv, epochs, batch, lr, pos_weight, seed, log, drop=()):
    n_param = rhz.Proposer.param_count()
    n_blob = rhz.Proposer.blob_count()
    mean = X.mean(0).cpu().numpy()
    std = X.std(0).cpu().numpy()
    istd = np.where(std > 1e-6, 1.0 / np.maximum(std, 1e-6), 1.0).astype(np.float32)
    for k in drop:  # a dropped feature normalises to 0 whatever its value: the network never sees it
        mean[k] = 0.0
        istd[k] = 0.0
    blob = init_blob(n_blob, n_param, seed, mean.astype(np.float32), istd)
    d_blob = torch.from_numpy(blob).cuda()
    d_m = torch.zeros(n_param, dtype=torch.float32, device="cuda")
    d_v = torch.zeros(n_param, dtype=torch.float32, device="cuda")
    d_g = torch.zeros(n_param, dtype=torch.float32, device="cuda")
    d_l = torch.zeros(1, dtype=torch.float32, device="cuda")
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    step = 0
    hist = []
    n = len(X)
    for ep in range(epochs):
        perm = torch.randperm(n, device="cuda", generator=gen)
        tot = 0.0
        for b0 in range(0, n - batch + 1, batch):
            idx = perm[b0:b0 + batch]
            xb = X[idx].contiguous()
            yb = Y[idx].contiguous()
            d_g.zero_()
            d_l.zero_()
            rhz.Proposer.grad(d_blob, xb, yb, batch, pos_weight, d_g, d_l)
            step += 1
            cur_lr = lr * 0.5 * (1.0 + math.cos(math.pi * ep / max(epochs, 1)))
            rhz.Proposer.adam(d_blob, d_g, d_m, d_v, n_param, cur_lr, step, grad_scale=1.0 / batch)
            tot += float(d_l.item())
        ev = evaluate(rhz, torch, d_blob, Xv, Yv)
        hist.append({"epoch": ep, "train_loss": tot / max(n // batch * batch, 1), **ev})
        log(f"epoch {ep}: loss {hist[-1]['train_loss']:.5f} val precision {ev['precision']:.4f} "
            f"recall {ev['recall']:.4f}")
    return d_blob.cpu().numpy(), hist


def write_inc(blob, path, summary):
    lines = [
        "// Built-in parameters of the sign-risk proposer (MLP 12-64-64-64-1), written by",
        "// rhizome_v24.15_amd/proposer_train.py: trained on the GPU with the library's own",
        f"// forward/backward/Adam kernels on {summary['train_meshes']} seeded random vessels",
        f"// ({summary['train_samples']} candidate voxels labelled by the exact winding kernel).",
        f"// Held-out ({summary['val_meshes']} meshes): precision {summary['val']['precision']:.4f}, "
        f"recall {summary['val']['recall']:.4f} at logit >= 0.",
        "// Layout: W1[12][64] b1 W2[64][64] b2 W3[64][64] b3 w4[64] b4 | mean[12] | 1/std[12].",
        f"static const float kProposerBlob[{len(blob)}] = {{",
    ]
    # 9 significant digits round-trip binary32 exactly
    vals = [f"{float(x):.8e}f" for x in blob.astype(np.float32)]
    for i in range(0, len(vals), 8):
        lines.append("    " + ", ".join(vals[i:i + 8]) + ",")
    lines.append("};")
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--meshes", type=int, default=200)
    ap.add_argument("--val-meshes", type=int, default=20)
    ap.add_argument("--epochs", type=int, default=12)
    ap.add_argument("--batch", type=int, default=32768)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--pos-weight", type=float, default=2.0)
    ap.add_argument("--seed", type=int, default=17)
    ap.add_argument("--out", default=HERE,
                    help="directory of the outputs: csrc/proposer_weights.inc (built-in variant) or "
                         "proposer_<variant>.npy, and proposer_[<variant>_]training.json")
    ap.add_argument("--variant", default="",
                    help="name of a non-built-in variant (its blob goes to proposer_<variant>.npy)")
    ap.add_argument("--drop-feature", type=int, action="append", default=[],
                    help="feature index the network must not see (repeatable)")
    args = ap.parse_args()

    def log(*a):
        print(*a, flush=True)

    import torch

    rhz, geom = _load()
    if not torch.cuda.is_available():
        raise SystemExit("training needs a GPU")
    t0 = time.time()
    train_seeds = [100_000 + args.seed * 1000 + i for i in range(args.meshes)]
    val_seeds = [900_000 + args.seed * 1000 + i for i in range(args.val_meshes)]
    X, Y, meta = generate(rhz, geom, torch, train_seeds, log)
    Xv, Yv, meta_v = generate(rhz, geom, torch, val_seeds, log)
    log(f"data: {len(X)} train / {len(Xv)} val samples, positive rate {float(Y.mean()):.4f}; "
        f"{time.time() - t0:.1f} s")
    blob, hist = train(rhz, torch, X, Y, Xv, Yv, args.epochs, args.batch, args.lr, args.pos_weight, args.seed, log,
                       drop=args.drop_feature)
    summary = {"train_meshes": len(meta), "val_meshes": len(meta_v), "train_samples": int(len(X)),
               "val_samples": int(len(Xv)), "epochs": args.epochs, "batch": args.batch, "lr": args.lr,
               "pos_weight": args.pos_weight, "seed": args.seed, "dropped_features": args.drop_feature,
               "val": hist[-1], "history": hist, "seconds": time.time() - t0,
               "train_tris_range": [min(m["tris"] for m in meta), max(m["tris"] for m in meta)]}
    os.makedirs(args.out, exist_ok=True)
    if args.variant:
        np.save(os.path.join(args.out, f"proposer_{args.variant}.npy"), blob.astype(np.float32))
        name = f"proposer_{args.variant}_training.json"
    else:
        os.makedirs(os.path.join(args.out, "csrc"), exist_ok=True)
        write_inc(blob, os.path.join(args.out, "csrc", "proposer_weights.inc"), summary)
        name = "proposer_training.json"
    with open(os.path.join(args.out, name), "w") as fh:
        json.dump(summary, fh, indent=1)
    log(json.dumps({k: v for k, v in summary.items() if k != "history"}))


if __name__ == "__main__":
    main()
