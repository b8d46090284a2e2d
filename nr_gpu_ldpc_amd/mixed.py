"""Mixed-batch workloads: codewords of both NR base graphs at every lifting size and rate.

A batch draws, per codeword, the base graph (BG1 or BG2 with equal probability), a lifting
size uniform over the 51 sizes of TS 38.212 Table 5.3.2-1, a nominal rate uniform in
[rate_lo, rate_hi] and an Eb/N0 uniform in [ebn0_lo, ebn0_hi].  The rate is realised by
keeping m parity rows of the base graph (rate = kb / (kb + m - 2) after puncturing the first
2Z columns):

    BG1 (kb = 22): m = clamp(round(22 / R - 20), 4, 46)   rates 0.917 .. 1/3
    BG2 (kb = 10): m = clamp(round(10 / R - 8), 4, 42)    rates 5/6 .. 1/5

so nominal rates outside a base graph's range are clamped to its nearest row count; the
realised rate of every codeword is returned with the batch.

Channel LLRs come from the library's Philox channel: each codeword is a full-rate codeword of
its (base graph, Z) whose first (kb + m) * Z decoder columns are kept (a prefix of a full-rate
codeword satisfies the first m parity rows), transmitted with the noise level of its own
realised rate.
"""
import numpy as np

import nr_gpu_ldpc_amd as nr

LIFTING_SIZES = sorted({a * (1 << j) for a in (2, 3, 5, 7, 9, 11, 13, 15) for j in range(8)
                        if a * (1 << j) <= 384})
assert len(LIFTING_SIZES) == 51

KB = {nr.BG1: 22, nr.BG2: 10}
ROWS = {nr.BG1: 46, nr.BG2: 42}
COLS = {nr.BG1: 68, nr.BG2: 52}


def rows_for_rate(kind, rate):
    kb = KB[kind]
    m = np.rint(kb / np.asarray(rate, dtype=np.float64) - (kb - 2)).astype(np.int64)
    return np.clip(m, 4, ROWS[kind])


def realised_rate(kind, m):
    return KB[kind] / (KB[kind] + np.asarray(m) - 2.0)


def draw_batch(n, seed, rate=(0.2, 8 / 9), ebn0=(0.0, 4.0), kinds=(nr.BG1, nr.BG2), sizes=None):
    """Random mixed batch: dict of numpy arrays kind, Z, m_rows, rate, ebn0 and the descriptor
    array (MIXED_CW_DTYPE) with codewords packed back to back in the LLR / HD buffers."""
    rng = np.random.default_rng(seed)
    sizes = LIFTING_SIZES if sizes is None else list(sizes)
    kind = np.asarray(kinds)[rng.integers(0, len(kinds), n)]
    Z = np.asarray(sizes)[rng.integers(0, len(sizes), n)]
    r_nom = rng.uniform(rate[0], rate[1], n)
    m = np.empty(n, dtype=np.int64)
    for k in kinds:
        sel = kind == k
        m[sel] = rows_for_rate(k, r_nom[sel])
    return describe(kind, Z, m, ebn0=rng.uniform(ebn0[0], ebn0[1], n))


def describe(kind, Z, m, ebn0=None):
    kind = np.asarray(kind, dtype=np.int64)
    Z = np.asarray(Z, dtype=np.int64)
    m = np.asarray(m, dtype=np.int64)
    kb = np.where(kind == nr.BG1, 22, 10)
    n_vn = (kb + m) * Z
    hd_bytes = (n_vn + 7) // 8
    d = np.zeros(len(kind), dtype=nr.MIXED_CW_DTYPE)
    d["kind"], d["Z"], d["m_rows"] = kind, Z, m
    d["llr_offset"] = np.concatenate([[0], np.cumsum(n_vn)[:-1]])
    d["hd_offset"] = np.concatenate([[0], np.cumsum(hd_bytes)[:-1]])
    rate = np.where(kind == nr.BG1, realised_rate(nr.BG1, m), realised_rate(nr.BG2, m))
    return {"kind": kind, "Z": Z, "m_rows": m, "n_vn": n_vn, "hd_bytes": hd_bytes, "rate": rate,
            "ebn0": ebn0, "info_bits": kb * Z, "desc": d,
            "llr_floats": int(n_vn.sum()), "hd_total": int(hd_bytes.sum())}


def make_channel(torch, batch, seed, device="cuda", contexts=None):
    """Channel LLRs [llr_floats] and transmitted bits [llr_floats] (bytes) of a drawn batch,
    generated on the GPU, one channel launch per (base graph, Z)."""
    llr = torch.zeros(batch["llr_floats"], dtype=torch.float32, device=device)
    bits = torch.zeros(batch["llr_floats"], dtype=torch.uint8, device=device)
    contexts = {} if contexts is None else contexts
    kind, Z, n_vn = batch["kind"], batch["Z"], batch["n_vn"]
    off = batch["desc"]["llr_offset"]
    for key in sorted(set(zip(kind.tolist(), Z.tolist()))):
        idx = np.nonzero((kind == key[0]) & (Z == key[1]))[0]
        code = contexts.get(key)
        if code is None:
            code = contexts[key] = nr.Code(key[0], key[1])
        sig = np.array([nr.noise_stddev(float(batch["ebn0"][i]), float(batch["rate"][i])) for i in idx],
                       dtype=np.float32)
        d_sig = torch.from_numpy(sig).to(device)
        tmp = torch.empty((len(idx), code.n_vn), dtype=torch.float32, device=device)
        tbits = torch.empty((len(idx), code.n_vn), dtype=torch.uint8, device=device)
        code.channel(seed, int(key[0]) << 40 | int(key[1]) << 24, len(idx), tmp.data_ptr(),
                     sigma_ptr=d_sig.data_ptr(), cw_bits_ptr=tbits.data_ptr())
        # keep each codeword's prefix, scattered to its own offset (index work on the device)
        lens = torch.from_numpy(n_vn[idx].astype(np.int64)).to(device)
        keep = torch.arange(code.n_vn, device=device)[None, :] < lens[:, None]
        rows = torch.repeat_interleave(torch.arange(len(idx), device=device), lens)
        k = torch.arange(rows.numel(), device=device) - (torch.cumsum(lens, 0) - lens)[rows]
        dst = torch.from_numpy(off[idx].astype(np.int64)).to(device)[rows] + k
        llr[dst] = tmp[keep]
        bits[dst] = tbits[keep]
    return llr, bits


def info_errors(torch, batch, hd, bits):
    """Information-bit errors of a decoded mixed batch, counted on the device.

    hd: packed hard decisions (uint8 tensor addressed by the descriptors' hd offsets);
    bits: transmitted bits (uint8 tensor addressed by the LLR offsets).  The first kb * Z
    columns of every codeword are compared (punctured information bits included).
    Returns (bit_errors, frame_errors, per-codeword error counts as a tensor)."""
    dev = hd.device
    n = len(batch["kind"])
    K = torch.from_numpy(np.asarray(batch["info_bits"], dtype=np.int64)).to(dev)
    hd_off = torch.from_numpy(batch["desc"]["hd_offset"].astype(np.int64)).to(dev)
    llr_off = torch.from_numpy(batch["desc"]["llr_offset"].astype(np.int64)).to(dev)
    cw = torch.repeat_interleave(torch.arange(n, device=dev), K)
    k = torch.arange(cw.numel(), device=dev) - (torch.cumsum(K, 0) - K)[cw]
    hb = 8 * hd_off[cw] + k
    got = (hd[hb >> 3].to(torch.int32) >> (hb & 7).to(torch.int32)) & 1
    err = (got != bits[llr_off[cw] + k].to(torch.int32)).to(torch.int64)
    per_cw = torch.zeros(n, dtype=torch.int64, device=dev).index_add_(0, cw, err)
    return int(per_cw.sum()), int((per_cw > 0).sum()), per_cw


def quantize(x, D=0.5, B=8):
    """The reference's quantize(D, x, B) (src/host_helpers.cpp:46-58) on the host: floor(x / D)
    clamped to [-2^(B-1), 2^(B-1) - 1], as int8 (the LDPC_LLR_I8 element type with scale D)."""
    q = np.floor(np.asarray(x, np.float64) / D)
    return np.clip(q, -(1 << (B - 1)), (1 << (B - 1)) - 1).astype(np.int8)
