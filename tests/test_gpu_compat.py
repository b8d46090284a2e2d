"""The reference-named host API on the GPU: the BER/FER driver (bin/ldpc_ber, built on
include/nr_gpu_ldpc.hpp) decodes what the oracle decodes, bit for bit, for BG1 Z=384 and for
the 12 x 24 Z=96 code of the reference's shipped example (T=5, SNR 0..10 dB)."""
import json
import os
import subprocess

import numpy as np
import pytest

from tests.oracle_binding import OR_BG1, OR_BG2, OR_WIFI

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "nr_gpu_ldpc_amd", "bin", "ldpc_ber")


def _run(args, tmp_path, dtype=np.float32):
    prefix = str(tmp_path / "dump")
    r = subprocess.run([DRIVER] + args + ["--dump", prefix], capture_output=True, text=True, timeout=100)
    assert r.returncode == 0, r.stderr
    lines = [json.loads(x) for x in r.stdout.splitlines()]
    llr = np.fromfile(prefix + ".llr", dtype=dtype)
    hd = np.fromfile(prefix + ".hd", dtype=np.uint8)
    return lines, llr, hd


def _check_against_oracle(oracle, code, llr, hd, n_vn, T):
    n_cw = llr.size // n_vn
    ohd, _, _ = oracle.decode(code, llr.reshape(n_cw, n_vn), T)
    want = np.packbits(ohd.reshape(-1), bitorder="little")
    assert np.array_equal(hd[:want.size], want)


def test_driver_bg1_z384(oracle, tmp_path):
    print("case driver BG1 Z=384", flush=True)
    args = ["--info-bits", "8424", "--rate", "0.3333333", "--iters", "5", "--snr-min", "1.0", "--snr-max", "1.5",
            "--snr-step", "0.5", "--codewords", "8", "--cw-per-mcw", "2", "--mcw", "4", "--streams", "2"]
    lines, llr, hd = _run(args, tmp_path)
    assert [x["snr_db"] for x in lines] == [1.0, 1.5]
    assert all(x["frames"] == 16 for x in lines)
    _check_against_oracle(oracle, oracle.code(OR_BG1, 384), llr, hd, 68 * 384, 5)


def test_driver_qc_z96_sweep(oracle, tmp_path):
    """Config of the reference's shipped example: 80 unique codewords x 8 streams per SNR."""
    print("case driver QC Z=96 SNR 0..10", flush=True)
    args = ["--code", "qc", "--Z", "96", "--iters", "5", "--snr-min", "0", "--snr-max", "10", "--snr-step", "1",
            "--codewords", "80", "--cw-per-mcw", "8", "--mcw", "10", "--streams", "8"]
    lines, llr, hd = _run(args, tmp_path)
    assert len(lines) == 11 and all(x["frames"] == 640 for x in lines)
    fer = [x["fer"] for x in lines]
    assert fer[0] > 0.5 and fer[-1] == 0.0
    assert all(a >= b - 0.05 for a, b in zip(fer, fer[1:]))
    _check_against_oracle(oracle, oracle.code(OR_WIFI, 96), llr, hd, 24 * 96, 5)


def test_driver_char_build_qc_z96_sweep(oracle, tmp_path):
    """The reference's default CHAR_PRC build through the same driver: int8 LLRs from
    quantize(0.5, x, 8), decoded bit for bit as the oracle's CHAR_PRC restatement decodes them."""
    print("case driver CHAR_PRC QC Z=96 SNR 0..10", flush=True)
    args = ["--code", "qc", "--Z", "96", "--iters", "5", "--snr-min", "0", "--snr-max", "10", "--snr-step", "1",
            "--codewords", "80", "--cw-per-mcw", "8", "--mcw", "10", "--streams", "8", "--prec", "char"]
    lines, q, hd = _run(args, tmp_path, dtype=np.int8)
    assert len(lines) == 11 and all(x["frames"] == 640 for x in lines)
    fer = [x["fer"] for x in lines]
    assert fer[0] > 0.5 and fer[-1] == 0.0
    n_vn = 24 * 96
    ohd, _, _, oor = oracle.decode_char(oracle.code(OR_WIFI, 96), q.reshape(-1, n_vn), 5)
    want = np.packbits(ohd.reshape(-1), bitorder="little")
    assert np.array_equal(hd[:want.size], want)
    assert oor == 0


def test_driver_half_build_qc_z96_sweep(oracle, tmp_path):
    """The reference's HALF_PRC build through the same driver (T = half): LLRs rounded to binary16
    as its reduce_precision<half> does (__float2half_rn, src/setup_exp.cu:33-38), decoded bit for
    bit as the oracle's HALF_PRC restatement decodes them."""
    print("case driver HALF_PRC QC Z=96 SNR 0..10", flush=True)
    args = ["--code", "qc", "--Z", "96", "--iters", "5", "--snr-min", "0", "--snr-max", "10", "--snr-step", "1",
            "--codewords", "80", "--cw-per-mcw", "8", "--mcw", "10", "--streams", "8", "--prec", "half"]
    lines, llr_h, hd = _run(args, tmp_path, dtype=np.uint16)
    assert len(lines) == 11 and all(x["frames"] == 640 for x in lines)
    fer = [x["fer"] for x in lines]
    assert fer[0] > 0.5 and fer[-1] == 0.0
    n_vn = 24 * 96
    ohd, _, _ = oracle.decode_half(oracle.code(OR_WIFI, 96), llr_h.reshape(-1, n_vn), 5)
    want = np.packbits(ohd.reshape(-1), bitorder="little")
    assert np.array_equal(hd[:want.size], want)


def test_driver_half_build_bg1_z384(oracle, tmp_path):
    """HALF_PRC build on the reference's shipped main() code (BG1 from setup_ldpc(8424, 1/3))."""
    print("case driver HALF_PRC BG1 Z=384", flush=True)
    args = ["--info-bits", "8424", "--rate", "0.3333333", "--iters", "5", "--snr-min", "1.0", "--snr-max", "1.5",
            "--snr-step", "0.5", "--codewords", "8", "--cw-per-mcw", "2", "--mcw", "4", "--streams", "2",
            "--prec", "half"]
    lines, llr_h, hd = _run(args, tmp_path, dtype=np.uint16)
    assert all(x["frames"] == 16 for x in lines)
    n_vn = 68 * 384
    ohd, _, _ = oracle.decode_half(oracle.code(OR_BG1, 384), llr_h.reshape(-1, n_vn), 5)
    want = np.packbits(ohd.reshape(-1), bitorder="little")
    assert np.array_equal(hd[:want.size], want)


def test_stream_packing_for_odd_lifting_size(nr, torch_gpu):
    """Per-codeword packed decisions -> one bit stream over [cw][n_vn] (n_vn % 8 != 0)."""
    torch = torch_gpu
    print("case stream pack BG2 Z=13", flush=True)
    code = nr.Code(2, 13)
    n = 37
    rng = np.random.default_rng(1)
    bits = rng.integers(0, 2, (n, code.n_vn), dtype=np.uint8)
    per_cw = np.packbits(bits, axis=1, bitorder="little")
    d_in = torch.from_numpy(per_cw).cuda()
    out_bytes = (n * code.n_vn + 7) // 8
    d_out = torch.zeros(out_bytes, dtype=torch.uint8, device="cuda")
    code.hd_to_stream(d_in.data_ptr(), n, d_out.data_ptr())
    torch.cuda.synchronize()
    assert code.n_vn % 8 != 0
    assert np.array_equal(d_out.cpu().numpy(), np.packbits(bits.reshape(-1), bitorder="little"))


C_ENTRY = os.path.join(ROOT, "tests", "compat", "_bin", "c_entry")


@pytest.mark.parametrize("A,rate,cpm,n_mcw,ebn0,bg,Z", [
    (8424, 1 / 3, 2, 4, 1.5, 1, 384),   # the reference's shipped main(): N_before_puncture % 8 == 0
    (62, 0.2, 3, 5, 2.0, 2, 13),        # odd Z: 52 * 13 = 676 columns, the stream-repack leg
])
def test_c_entry_point_matches_oracle(oracle, tmp_path, A, rate, cpm, n_mcw, ebn0, bg, Z):
    """ldpc_gpu_decoder(char*, char*) (Linking_With_Other_Projects.md:19-25) on the int8 LLRs that
    quantize(0.5, x, 8) makes: its packed output stream equals the oracle's CHAR_PRC decode of
    the same inputs, packed over the concatenated [n_cw][N_before_puncture] columns as
    pack_hard_decision lays them out; and it equals the C ABI's decode + stream repack."""
    print(f"case C entry point A={A} BG{bg} Z={Z}", flush=True)
    prefix = str(tmp_path / "centry")
    n_iter = 10
    r = subprocess.run([C_ENTRY, str(A), str(rate), str(cpm), str(n_mcw), str(n_iter), str(ebn0), prefix],
                       capture_output=True, text=True, timeout=100)
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout)
    n_cw = cpm * n_mcw
    n_vn = out["n_vn"]
    assert (out["bg"], out["Z"], out["n_cw"]) == (bg, Z, n_cw)
    assert n_vn == (68 if bg == 1 else 52) * Z
    assert out["status"] == 0 and out["differing_bytes"] == 0 and out["bytes"] == (n_cw * n_vn + 7) // 8
    q = np.fromfile(prefix + ".llr", dtype=np.int8).reshape(n_cw, n_vn)
    got = np.fromfile(prefix + ".out", dtype=np.uint8)
    ohd, _, _, oor = oracle.decode_char(oracle.code(OR_BG1 if bg == 1 else OR_BG2, Z), q, n_iter)
    want = np.packbits(ohd.reshape(-1), bitorder="little")
    assert np.array_equal(got, want)
    assert 0 < int(np.unpackbits(got).sum()) < n_cw * n_vn  # neither all zeros nor all ones
    assert oor == 0
