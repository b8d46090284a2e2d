"""Training and evaluation of the learned offset/normalized min-sum decoder on the GPU.

All arithmetic runs in the library's HIP kernels: the Philox channel generates the training
and test codewords, the learned forward (with training records), the hand-written backward
and the Adam update.  PyTorch only provides device buffers and the stream.

    python -m nr_gpu_ldpc_amd.noms train --kind 2 --Z 208 --iters 10 --codewords 1048576
    python -m nr_gpu_ldpc_amd.noms eval  --kind 2 --Z 208 --iters 10

Weights are stored as versioned binary files (see save_weights) under nr_gpu_ldpc_amd/weights/.
"""
import argparse
import json
import math
import os
import struct
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import nr_gpu_ldpc_amd as nr  # noqa: E402

WEIGHTS_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights")
MAGIC = b"NRLDPCW1"


def weights_path(kind, Z, n_iter):
    return os.path.join(WEIGHTS_DIR, f"noms_k{kind}_z{Z}_t{n_iter}.bin")


def save_weights(path, kind, Z, n_iter, w, meta=None):
    """Layout: magic(8) | u32 version=1 | i32 kind, Z, n_iter, nnz_full | u32 meta_len |
    meta JSON (utf-8) | float32 weights [n_iter][nnz_full][2] little endian."""
    os.makedirs(os.path.dirname(path), exist_ok=True)
    w = np.ascontiguousarray(w, dtype="<f4")
    m = json.dumps(meta or {}, sort_keys=True).encode()
    with open(path, "wb") as f:
        f.write(MAGIC + struct.pack("<Iiiii", 1, kind, Z, n_iter, w.shape[1]) + struct.pack("<I", len(m)) + m)
        f.write(w.tobytes())


def load_weights(path):
    with open(path, "rb") as f:
        head = f.read(8 + 20)
        if head[:8] != MAGIC:
            raise ValueError(f"{path}: not a weights file")
        version, kind, Z, n_iter, nnz = struct.unpack("<Iiiii", head[8:])
        if version != 1:
            raise ValueError(f"{path}: unsupported version {version}")
        (mlen,) = struct.unpack("<I", f.read(4))
        meta = json.loads(f.read(mlen).decode() or "{}")
        w = np.frombuffer(f.read(), dtype="<f4").reshape(n_iter, nnz, 2).copy()
    return {"kind": kind, "Z": Z, "n_iter": n_iter, "weights": w, "meta": meta}


def sigmas_for(ebn0, rate):
    return np.array([nr.noise_stddev(float(e), rate) for e in ebn0], dtype=np.float32)


def code_rate(code):
    return code.K / code.N if code.punct else code.kb / code.cols


def degree1_edges(kind, Z):
    """Boolean mask over the base graph's edges (row-major, all rows) that end in a degree-1
    column.  Their messages only set the posterior of that column (its single check receives
    the channel LLR), so they shape the output of the extension parity bits and nothing else."""
    code = nr.Code(kind, Z, 0, -1)
    h = nr.base_matrix(kind, code.ils)
    deg = (h >= 0).sum(axis=0)
    return np.array([deg[j] == 1 for i in range(h.shape[0]) for j in range(h.shape[1]) if h[i, j] >= 0])


def loss_weight_fn(spec):
    """Per-frame loss weight as a function of the frame's Eb/N0 (dB), from a text spec:
    "none" (all 1), "sigmoid:c,w" (1 / (1 + exp(-(e - c) / w)): frames well below the decoder's
    threshold, whose BCE is large and whose bits no weight choice will fix, count little) or
    "cutoff:c" (1 at e >= c, else 0)."""
    if not spec or spec == "none":
        return None
    name, _, args = spec.partition(":")
    vals = [float(v) for v in args.split(",") if v]
    if name == "sigmoid" and len(vals) == 2:
        c, w = vals
        return lambda e: 1.0 / (1.0 + np.exp(-(np.asarray(e, np.float64) - c) / w))
    if name == "cutoff" and len(vals) == 1:
        return lambda e: (np.asarray(e, np.float64) >= vals[0]).astype(np.float64)
    raise ValueError(f"unknown loss weight spec {spec!r}")


def train(kind=2, Z=208, n_iter=10, codewords=1 << 20, batch=1024, lr=2e-3, ebn0=(-1.0, 3.0), seed=1,
          device=0, log=print, freeze_degree1=False, loss_weight=None, epochs=1):
    """loss_weight: None, or a function of the frames' Eb/N0 array giving each frame's loss weight
    (loss_weight_fn).  epochs: passes over the same `codewords` training codewords (the channel
    regenerates them bit for bit from the seed, each with its own Eb/N0)."""
    import torch
    code = nr.Code(kind, Z, 0, device)
    code.set_weights(code.plain_weights(n_iter))
    dev = torch.device(f"cuda:{device}")
    s = torch.cuda.current_stream(dev).cuda_stream
    llr = torch.empty((batch, code.n_vn), dtype=torch.float32, device=dev)
    bits = torch.empty((batch, code.n_vn), dtype=torch.uint8, device=dev)
    sig = torch.empty(batch, dtype=torch.float32, device=dev)
    shape = (n_iter, code.nnz_full, 2)
    grad = torch.zeros(shape, dtype=torch.float32, device=dev)
    mom = torch.zeros(shape, dtype=torch.float32, device=dev)
    vel = torch.zeros(shape, dtype=torch.float32, device=dev)
    loss = torch.zeros(1, dtype=torch.float32, device=dev)
    wts = torch.ones(batch, dtype=torch.float32, device=dev)
    keep = None
    if freeze_degree1:
        keep = torch.from_numpy((~degree1_edges(kind, Z)).astype(np.float32)).to(dev).view(1, -1, 1)
    ws = torch.empty(code.workspace_bytes(batch, n_iter), dtype=torch.uint8, device=dev)
    rate = code_rate(code)
    per_epoch = max(1, codewords // batch)
    steps = per_epoch * max(1, epochs)
    hist = []
    t0 = time.time()
    rng = None
    for step in range(1, steps + 1):
        if (step - 1) % per_epoch == 0:
            rng = np.random.default_rng(seed)  # each epoch draws the same Eb/N0 sequence
        eb = rng.uniform(ebn0[0], ebn0[1], batch)
        sig.copy_(torch.from_numpy(sigmas_for(eb, rate)))
        if loss_weight is not None:
            wts.copy_(torch.from_numpy(np.asarray(loss_weight(eb), dtype=np.float32)))
        code.channel(seed, ((step - 1) % per_epoch) * batch, batch, llr.data_ptr(), sigma_ptr=sig.data_ptr(),
                     cw_bits_ptr=bits.data_ptr(), stream=s)
        grad.zero_()
        loss.zero_()
        code.train_grad(llr.data_ptr(), bits.data_ptr(), batch, n_iter, grad.data_ptr(), loss.data_ptr(),
                        ws.data_ptr(), ws.numel(), stream=s, weight_ptr=wts.data_ptr() if loss_weight else 0)
        if keep is not None:
            grad.mul_(keep)
        step_lr = lr * 0.5 * (1 + math.cos(math.pi * (step - 1) / steps))  # cosine decay
        code.adam(grad.data_ptr(), mom.data_ptr(), vel.data_ptr(), step, step_lr, stream=s)
        if step % max(1, steps // 20) == 0 or step == steps:
            torch.cuda.synchronize(dev)
            hist.append((step, float(loss.item())))
            log(f"step {step}/{steps} loss {hist[-1][1]:.5f} ({time.time() - t0:.1f} s)")
    torch.cuda.synchronize(dev)
    return code, code.get_weights(), hist, time.time() - t0


def fer_curve(code, n_iter, ebn0_list, n_cw, learned=False, batch=8192, seed=99, flags=0, device=0):
    """Frame/bit error rates over Eb/N0 with fresh Philox codewords (decode on the GPU)."""
    import torch
    dev = torch.device(f"cuda:{device}")
    s = torch.cuda.current_stream(dev).cuda_stream
    llr = torch.empty((batch, code.n_vn), dtype=torch.float32, device=dev)
    bits = torch.empty((batch, code.n_vn), dtype=torch.uint8, device=dev)
    hd = torch.empty((batch, code.hd_bytes), dtype=torch.uint8, device=dev)
    cnt = torch.zeros(2, dtype=torch.int64, device=dev)
    wgt = torch.empty(batch, dtype=torch.int32, device=dev)
    its = torch.empty(batch, dtype=torch.int32, device=dev)
    rate = code_rate(code)
    out = []
    for pi, e in enumerate(ebn0_list):
        sigma = nr.noise_stddev(float(e), rate)
        cnt.zero_()
        synd_fail = torch.zeros((), dtype=torch.int64, device=dev)
        it_sum = torch.zeros((), dtype=torch.int64, device=dev)
        done = 0
        while done < n_cw:
            b = min(batch, n_cw - done)
            code.channel(seed + pi, done, b, llr.data_ptr(), sigma=sigma, cw_bits_ptr=bits.data_ptr(), stream=s)
            if learned:
                code.decode_learned(llr.data_ptr(), b, n_iter, hd.data_ptr(), 0, its.data_ptr(), flags=flags, stream=s)
            else:
                code.decode(llr.data_ptr(), b, n_iter, hd.data_ptr(), 0, its.data_ptr(), flags=flags, stream=s)
            code.count_errors(hd.data_ptr(), bits.data_ptr(), b, cnt.data_ptr(), stream=s)
            code.syndrome(hd.data_ptr(), b, wgt.data_ptr(), stream=s)
            synd_fail += torch.count_nonzero(wgt[:b])
            it_sum += its[:b].sum()
            done += b
        torch.cuda.synchronize(dev)
        be, fe = cnt.tolist()
        out.append({"ebn0_db": float(e), "fer": fe / n_cw, "ber": be / (n_cw * code.K), "frames": n_cw,
                    "frame_errors": fe, "syndrome_fail": int(synd_fail) / n_cw,
                    "mean_iters": int(it_sum) / n_cw})
    return out


def crossing(curve, target=1e-3):
    """Eb/N0 where the FER curve crosses `target` (log-linear interpolation), or None."""
    pts = [(p["ebn0_db"], p["fer"]) for p in curve]
    for (e0, f0), (e1, f1) in zip(pts, pts[1:]):
        if f0 >= target > f1:
            if f1 <= 0:
                return e1
            l0, l1, lt = math.log10(f0), math.log10(f1), math.log10(target)
            return e0 + (e1 - e0) * (l0 - lt) / (l0 - l1)
    return None


def main():
    p = argparse.ArgumentParser()
    p.add_argument("mode", choices=["train", "eval"])
    p.add_argument("--kind", type=int, default=2)
    p.add_argument("--Z", type=int, default=208)
    p.add_argument("--iters", type=int, default=10)
    p.add_argument("--codewords", type=int, default=1 << 20)
    p.add_argument("--batch", type=int, default=1024)
    p.add_argument("--lr", type=float, default=2e-3)
    p.add_argument("--eval-codewords", type=int, default=200000)
    p.add_argument("--ebn0", type=str, default="0.5,0.75,1.0,1.25,1.5,1.75,2.0")
    p.add_argument("--out", type=str, default="")
    p.add_argument("--ebn0-train", type=str, default="-1,3", help="training Eb/N0 range lo,hi (dB)")
    p.add_argument("--seed", type=int, default=1)
    p.add_argument("--plain-iters", type=int, default=0, help="plain decoder iterations in eval (default: --iters)")
    p.add_argument("--early-exit", action="store_true", help="evaluate with syndrome early exit")
    p.add_argument("--freeze-degree1", action="store_true",
                   help="keep the weights of edges into degree-1 columns at their plain values")
    p.add_argument("--epochs", type=int, default=1, help="passes over the same training codewords")
    p.add_argument("--loss-weight", type=str, default="none",
                   help='per-frame loss weight by Eb/N0: "none", "sigmoid:c,w" or "cutoff:c"')
    a = p.parse_args()
    path = a.out or weights_path(a.kind, a.Z, a.iters)
    if a.mode == "train":
        rng_db = tuple(float(x) for x in a.ebn0_train.split(","))
        code, w, hist, secs = train(a.kind, a.Z, a.iters, a.codewords, a.batch, a.lr, rng_db, a.seed,
                                   freeze_degree1=a.freeze_degree1, loss_weight=loss_weight_fn(a.loss_weight),
                                   epochs=a.epochs)
        save_weights(path, a.kind, a.Z, a.iters, w, {"codewords": a.codewords, "batch": a.batch, "lr": a.lr,
                                                     "ebn0_train": list(rng_db), "final_loss": hist[-1][1],
                                                     "train_seconds": round(secs, 1), "seed": a.seed,
                                                     "freeze_degree1": a.freeze_degree1,
                                                     "loss_weight": a.loss_weight, "epochs": a.epochs})
        print(json.dumps({"saved": os.path.relpath(path, ROOT), "loss": hist, "seconds": round(secs, 1)}))
    else:
        rec = load_weights(path)
        code = nr.Code(a.kind, a.Z)
        code.set_weights(rec["weights"])
        ebs = [float(x) for x in a.ebn0.split(",")]
        ee = nr.EARLY_EXIT if a.early_exit else 0
        plain = fer_curve(code, a.plain_iters or a.iters, ebs, a.eval_codewords, flags=ee)
        learned = fer_curve(code, a.iters, ebs, a.eval_codewords, learned=True, flags=ee)
        fallback = fer_curve(code, a.iters, ebs, a.eval_codewords, learned=True, flags=ee | nr.FALLBACK_PLAIN)
        print(json.dumps({"plain": plain, "learned": learned, "learned_fallback": fallback,
                          "plain_fer1e-3_db": crossing(plain), "learned_fer1e-3_db": crossing(learned),
                          "learned_fallback_fer1e-3_db": crossing(fallback), "meta": rec["meta"]}))


if __name__ == "__main__":
    main()
