"""Learned entry-leaf predictor: training labels from the plain search, GPU
training of the MLP (librte kernels), and inference with the kd-tree leaf as
fallback.

Labels.  For a training query every leaf is a candidate entry (or a given
shortlist, e.g. the kd-tree leaf and the leaves of the exact 10 nearest
neighbours).  Each candidate seeds the plain search under a ladder of
iteration budgets; its score is the number of exact top-10 neighbours found,
summed over the budgets -- the area under its recall-versus-hops curve -- so
the label is the entry that converges in the fewest hops.  Ties keep the
kd-tree leaf, then the lowest leaf id.
"""
from __future__ import annotations

import numpy as np

from .api import Context, make_params


def leaf_of_points(pids: np.ndarray, n: int) -> np.ndarray:
    """belong[i] = leaf holding point i (from the ascending pid table)."""
    belong = np.full(n, -1, np.int32)
    leaf = np.repeat(np.arange(pids.shape[0], dtype=np.int32), pids.shape[1])
    flat = pids.reshape(-1)
    ok = flat < n
    belong[flat[ok]] = leaf[ok]
    return belong


def candidate_leaves(kd_hits: np.ndarray, gt: np.ndarray, belong: np.ndarray, k: int = 10):
    """(nq, 1 + k) candidate leaves, -1 padded: kd leaf then the distinct
    leaves of the k exact neighbours."""
    nq = kd_hits.shape[0]
    nbr = belong[gt[:, :k]]
    cand = np.full((nq, 1 + k), -1, np.int32)
    cand[:, 0] = kd_hits
    for q in range(nq):
        seen = {int(kd_hits[q])}
        c = 1
        for x in nbr[q]:
            x = int(x)
            if x >= 0 and x not in seen:
                seen.add(x)
                cand[q, c] = x
                c += 1
    return cand


def make_labels(ctx: Context, PQ, cand, gt, beam: int = 64, width: int = 1,
                budgets=(4, 8, 16), mode: int = 0, kd=None, n_leaf: int | None = None,
                rows_per_call: int = 1 << 22, top: int = 1, log=None):
    """Label per query (device search, device scoring).

    cand: (nq, C) int32 shortlist, -1 padded, column 0 the kd-tree leaf; or
    None to score all n_leaf leaves (kd = the kd-tree leaves for tie-breaks).
    Returns (labels int32[nq], best score int32[nq]); with top > 1 (all-leaf
    scoring only) labels is (nq, top), the best `top` leaves in key order.
    log: optional callable for progress lines (all-leaf scoring)."""
    import torch
    dev = PQ.device
    nq = PQ.shape[0]
    gt10 = torch.as_tensor(np.ascontiguousarray(gt[:, :10]), device=dev).int()
    params = [make_params(k=10, beam=beam, width=width, max_iter=b, mode=mode) for b in budgets]
    if cand is None:
        L = int(n_leaf)
        kd_t = torch.as_tensor(np.ascontiguousarray(kd), device=dev).long()
        labels = np.empty((nq, top) if top > 1 else nq, np.int32)
        best = np.empty(nq, np.int32)
        qc = max(1, rows_per_call // L)
        leaf_ids = torch.arange(L, device=dev, dtype=torch.int32)
        for q0 in range(0, nq, qc):
            qn = min(qc, nq - q0)
            qi = torch.arange(q0, q0 + qn, device=dev).repeat_interleave(L)
            X = PQ[qi].contiguous()
            leaf = leaf_ids.repeat(qn).contiguous()
            g = gt10[qi]
            hits = torch.zeros(qn * L, dtype=torch.int64, device=dev)
            for p in params:
                ids = ctx.search(X, leaf, p, with_dists=False)
                hits += (ids[:, :, None] == g[:, None, :]).any(-1).sum(1)
            hits = hits.view(qn, L)
            # unique key: hits, then the kd-tree leaf, then the lowest id
            key = hits * (2 * L) + (L - 1 - leaf_ids.long())[None, :]
            key[torch.arange(qn, device=dev), kd_t[q0:q0 + qn]] += L
            arg = key.argmax(1)
            if top > 1:
                labels[q0:q0 + qn] = key.topk(top, dim=1).indices.int().cpu().numpy()
            else:
                labels[q0:q0 + qn] = arg.int().cpu().numpy()
            best[q0:q0 + qn] = hits.gather(1, arg[:, None])[:, 0].int().cpu().numpy()
            del X, g, hits, key
            if log is not None and (q0 // qc) % 16 == 15:
                log(f"labels {q0 + qn}/{nq} queries")
        return labels, best
    C = cand.shape[1]
    qi_np, ci_np = np.nonzero(cand >= 0)
    leaves = torch.from_numpy(cand[qi_np, ci_np].astype(np.int32)).to(dev)
    qidx = torch.from_numpy(qi_np.astype(np.int64)).to(dev)
    hits = torch.zeros(qi_np.size, dtype=torch.int64, device=dev)
    for p in params:
        for s in range(0, qi_np.size, rows_per_call):
            sl = slice(s, min(qi_np.size, s + rows_per_call))
            ids = ctx.search(PQ[qidx[sl]].contiguous(), leaves[sl].contiguous(), p,
                             with_dists=False)
            hits[sl] += (ids[:, :, None] == gt10[qidx[sl]][:, None, :]).any(-1).sum(1)
    h = hits.cpu().numpy()
    order = np.lexsort((ci_np, -h, qi_np))
    first = np.ones(order.size, bool)
    first[1:] = qi_np[order][1:] != qi_np[order][:-1]
    top = order[first]
    labels = np.empty(nq, np.int32)
    best = np.empty(nq, np.int32)
    labels[qi_np[top]] = cand[qi_np[top], ci_np[top]]
    best[qi_np[top]] = h[top]
    return labels, best


def train(ctx: Context, X, y, n_out: int, hidden=(256, 256), epochs: int = 20,
          batch: int = 1024, lr: float = 2e-3, seed: int = 1, log=None, validate=None,
          every: int = 2, input_norm: str = "feature"):
    """Trains the MLP on device tensors X (n x d) and labels y (n).  The
    input is centred with the training set's mean and scaled (kept in the
    network): input_norm "feature" divides every feature by its own standard
    deviation; "global" divides all features by one scale, the root mean
    variance, which keeps the relative variances (for PCA coordinates whose
    trailing columns are mostly noise, per-feature scaling would lift that
    noise to the weight of the leading columns).

    validate: optional callable returning a score (higher is better) of the
    current network, e.g. end-to-end recall with the kd-tree fallback on held
    out queries; it runs every `every` epochs and the best weights are kept
    (early stopping on the quantity that matters, not on the training loss).
    Returns (losses per epoch, validation history [(epoch, score)])."""
    import torch
    n, d = X.shape
    ctx.mlp_create(d, hidden[0], hidden[1], n_out, seed=seed)
    Xh = X.double()
    mean = Xh.mean(0)
    if input_norm == "feature":
        std = Xh.std(0).clamp_min(1e-6)
    elif input_norm == "global":
        std = Xh.var(0).mean().sqrt().clamp_min(1e-6).repeat(d)
    else:
        raise ValueError(f"input_norm {input_norm!r}")
    ctx.mlp_set_input_norm(mean.float().cpu().numpy(), (1.0 / std).float().cpu().numpy())
    g = torch.Generator(device="cpu")
    g.manual_seed(seed)
    losses, history = [], []
    best = None
    for ep in range(epochs):
        perm = torch.randperm(n, generator=g).to(X.device)
        Xs, ys = X[perm].contiguous(), y[perm].contiguous()
        last = None
        for s in range(0, n - batch + 1, batch):
            want = log is not None or (s + 2 * batch > n)
            l = ctx.mlp_train_step(Xs[s:s + batch], ys[s:s + batch], lr=lr, want_loss=want)
            if l is not None:
                last = l
        losses.append(last)
        if validate is not None and ((ep + 1) % every == 0 or ep == epochs - 1):
            score = float(validate())
            history.append((ep + 1, score))
            if best is None or score > best[0]:
                best = (score, ep + 1, ctx.mlp_get())
            if log:
                log(f"epoch {ep} loss {last:.4f} validation {score:.4f}")
        elif log:
            log(f"epoch {ep} loss {last:.4f}")
    if best is not None:
        ctx.mlp_set(*best[2])
    return losses, history


def predict(ctx: Context, X, fallback, min_conf: float = 0.5, conf=None):
    return ctx.mlp_predict(X, fallback=fallback, min_conf=min_conf, conf=conf)


def scale_output(ctx: Context, factor: float):
    """Multiply the last layer's weights and bias by `factor` (a softmax
    temperature of 1/factor): the arg-max leaf is unchanged up to bf16
    rounding, only the confidence used by the fallback rule moves."""
    W, b = ctx.mlp_get()
    W[2] = W[2] * np.float32(factor)
    b[2] = b[2] * np.float32(factor)
    ctx.mlp_set(W, b)


def fit_temperature(ctx: Context, validate,
                    factors=(1.0, 1.5, 2.0, 3.0, 4.0, 6.0, 8.0, 12.0, 16.0)):
    """Choose the output scale (inverse softmax temperature) that maximises
    `validate()` -- end-to-end recall on held-out queries with the fallback
    rule applied -- and leave the network scaled by it.  Returns (factor,
    [(factor, score)])."""
    W0, b0 = ctx.mlp_get()
    hist = []
    for f in factors:
        W = [w.copy() for w in W0]
        b = [x.copy() for x in b0]
        W[2] *= np.float32(f)
        b[2] *= np.float32(f)
        ctx.mlp_set(W, b)
        hist.append((f, float(validate())))
    best = max(hist, key=lambda x: (x[1], -x[0]))
    W = [w.copy() for w in W0]
    b = [x.copy() for x in b0]
    W[2] *= np.float32(best[0])
    b[2] *= np.float32(best[0])
    ctx.mlp_set(W, b)
    return best[0], hist


def save(ctx: Context, path: str, **extra):
    """Write the trained network (weights, biases, input normalisation) and
    `extra` scalars to an .npz file."""
    W, b = ctx.mlp_get()
    d = {f"W{i}": W[i] for i in range(3)}
    d.update({f"b{i}": b[i] for i in range(3)})
    norm = getattr(ctx, "mlp_norm", None)
    if norm is not None:
        d["norm_mean"], d["norm_inv"] = norm
    d.update({k: np.asarray(v) for k, v in extra.items()})
    np.savez(path, **d)


def load(ctx: Context, path: str) -> dict:
    """Create the network on `ctx` from a file written by save(); returns
    the extra scalars."""
    z = np.load(path)
    W = [z[f"W{i}"] for i in range(3)]
    b = [z[f"b{i}"] for i in range(3)]
    ctx.mlp_create(W[0].shape[0], W[0].shape[1], W[1].shape[1], W[2].shape[1])
    ctx.mlp_set(W, b)
    if "norm_mean" in z:
        ctx.mlp_set_input_norm(z["norm_mean"], z["norm_inv"])
    skip = {f"W{i}" for i in range(3)} | {f"b{i}" for i in range(3)} | {"norm_mean", "norm_inv"}
    return {k: z[k].item() for k in z.files if k not in skip}
