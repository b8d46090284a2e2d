"""The reference-named C++ host API (include/nr_gpu_ldpc.hpp), host side: a C++ program built
against the library prints its results, checked here against the C ABI, the oracle and
closed forms."""
import ctypes
import json
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "nr_gpu_ldpc_amd", "lib")


@pytest.fixture(scope="module")
def host_out(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("compat") / "compat_host")
    cmd = ["/opt/rocm/bin/hipcc", "-O1", "-std=c++17", f"-I{ROOT}/include", f"{ROOT}/tests/compat/compat_host.cpp",
           f"-L{LIBDIR}", "-lnrldpc", f"-Wl,-rpath,{LIBDIR}", "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True)
    r = subprocess.run([exe], check=True, capture_output=True, text=True)
    return json.loads(r.stdout)


def test_compact_lists(host_out, nr):
    """fill_in_compact_matrices: one entry per nonzero block, row lists in ascending column,
    column lists in ascending row, unused slots {0, 0, -1, 0}."""
    H = np.array(host_out["H_bg1_row0"])
    n1, n2, pad_v, pad_valid, pad2_v = host_out["compact_counts"]
    assert n1 == n2 == 316
    assert pad_v == -1 and pad_valid == 0 and pad2_v == -1
    row0 = np.array(host_out["compact_row0"]).reshape(19, 3)
    cols = np.nonzero(H >= 0)[0]
    assert (row0[:, 0] == 0).all() and list(row0[:, 1]) == list(cols) and list(row0[:, 2]) == list(H[cols])
    col0 = np.array(host_out["compact_col0"]).reshape(30, 3)
    base = nr.base_matrix(nr.BG1, 1)  # Z = 384 is in lifting set 1
    rows = np.nonzero(base[:, 0] >= 0)[0]
    assert (col0[:, 1] == 0).all()
    assert list(col0[:, 0]) == list(rows) and list(col0[:, 2]) == list(base[rows, 0])


def test_allocate_constant_memory_refuses_a_wrong_list(tmp_path):
    """A row list that does not describe ldpc->H is refused (message, abort) before any device
    call, so this runs without a GPU."""
    exe = str(tmp_path / "compat_host")
    subprocess.run(["/opt/rocm/bin/hipcc", "-O1", "-std=c++17", f"-I{ROOT}/include",
                    f"{ROOT}/tests/compat/compat_host.cpp", f"-L{LIBDIR}", "-lnrldpc", f"-Wl,-rpath,{LIBDIR}",
                    "-o", exe], check=True, capture_output=True)
    r = subprocess.run([exe, "corrupt"], capture_output=True, text=True)
    assert r.returncode != 0
    assert "h_compact1 does not describe ldpc->H" in r.stderr


def test_setup_host_refuses_a_batch_beyond_the_int_size_fields(tmp_path):
    """host_mem keeps per-stream byte counts in int fields; a batch whose LLR bytes exceed
    INT_MAX is refused (message, abort) before any pinned allocation, so this runs without a GPU."""
    exe = str(tmp_path / "compat_host")
    subprocess.run(["/opt/rocm/bin/hipcc", "-O1", "-std=c++17", f"-I{ROOT}/include",
                    f"{ROOT}/tests/compat/compat_host.cpp", f"-L{LIBDIR}", "-lnrldpc", f"-Wl,-rpath,{LIBDIR}",
                    "-o", exe], check=True, capture_output=True)
    r = subprocess.run([exe, "huge"], capture_output=True, text=True)
    assert r.returncode != 0
    assert "LLR bytes fit an int" in r.stderr


def test_setup_ldpc(host_out):
    # BG1 Z=384: K=8448, N=25344, 46 x 68, 26112 decoder columns, row degree 19, column degree 30
    assert host_out["setup_bg1"] == [1, 384, 8448, 25344, 46, 68, 26112, 19, 30, 5, 19, 30]
    assert host_out["setup_bg2"][:7] == [2, 208, 2080, 10400, 42, 52, 10816]
    assert host_out["setup_qc"] == [3, 96, 1152, 2304, 12, 24, 2304]


def test_base_matrix_row(host_out, nr):
    assert host_out["H_bg1_row0"] == nr.base_matrix(1, 1)[0].tolist()


def _full_codeword(nr, kind, Z, info, out, kb, n_par):
    """Rebuild [all info | parity] from the reference output layout and check H c = 0."""
    code = nr.Code(kind, Z, 0, -1)
    want = code.encode_host(np.array(info + [0] * (kb * Z - len(info)), dtype=np.uint8)[None])[0]
    got_info = np.array(info[2 * Z:])
    assert np.array_equal(np.array(out[:len(got_info)]), got_info)
    assert np.array_equal(np.array(out[len(got_info):len(got_info) + n_par]), want[kb * Z:kb * Z + n_par])


def test_encoder_orig_layout(host_out, nr):
    assert host_out["enc_bg1_rc"] == 0 and host_out["enc_gen_code_rc"] == -1
    # BG1, 44 bits -> Z = 2: out = info bits 4..43 then all 46 x 2 parity bits
    _full_codeword(nr, 1, 2, host_out["enc_bg1_info"], host_out["enc_bg1_out"], 22, 92)
    # BG2, 40 bits: Kb = 6 -> Z = 7 (>= 40/6); 28 punctured parity columns and 4 removed bits
    # (src/nr_ldpc.cpp:736-737) leave (42 - 28) * 7 - 4 = 94 parity bits after the 26 info bits
    assert host_out["enc_bg2_rc"] == 0
    _full_codeword(nr, 2, 7, host_out["enc_bg2_info"], host_out["enc_bg2_out"], 10, 94)


def test_channel_helpers(host_out, nr, oracle):
    f32 = np.float32  # the program prints float32 values with 9 significant digits (exact round trip)
    assert f32(host_out["sigma"][0]) == f32(nr.noise_stddev(1.5, 1 / 3))
    assert f32(host_out["sigma"][1]) == f32(nr.noise_stddev(-2.0, 0.5))
    s = np.float32(0.8)
    recv = np.array([0.5, -1.25, 2.0, 0.0, -0.75, 1.0], np.float32)
    want = [0.0, 0.0] + [float((r * np.float32(2)) / (s * s)) for r in recv]
    assert np.array_equal(np.array(host_out["llr"], np.float32), np.array(want, np.float32))
    assert host_out["error_check"] == 2
    assert host_out["cuda_error_check"] == [3, 6, 1, 2]  # codeword 0 clean, codeword 1: 3 errors
    assert host_out["quantize"] == [2, -128, 127, -1]
    assert host_out["unpack"] == [1, 0, 1, 0, 0, 1, 0, 1, 1, 1, 0, 0]


def test_transmit_reproduces_the_reference_random_stream(host_out, nr):
    """Same libc calls: srand/rand for the bits, srand48/drand48 polar draws for the noise."""
    libc = ctypes.CDLL(None)
    libc.rand.restype = ctypes.c_int
    libc.drand48.restype = ctypes.c_double
    libc.srand(4242)
    libc.srand48(99)
    K, nb = 44, 136
    info = [libc.rand() % 2 for _ in range(K)]
    assert host_out["tx_info"] == info
    code = nr.Code(1, 2, 0, -1)
    cw = code.encode_host(np.array(info, np.uint8)[None])[0]
    want_cw = np.zeros(nb, int)
    want_cw[:K - 4] = info[4:]
    want_cw[K - 4:K - 4 + 92] = cw[44:136]
    assert host_out["tx_cw"] == want_cw.tolist()
    x = np.where(want_cw == 0, np.float32(1), np.float32(-1)).astype(np.float32)
    recv = []
    for i in range(nb):
        while True:
            u1 = np.float32(np.float32(libc.drand48()) * np.float32(2) - np.float32(1))
            u2 = np.float32(np.float32(libc.drand48()) * np.float32(2) - np.float32(1))
            s = np.float32(u1 * u1 + u2 * u2)
            if s < 1:
                break
        noise = np.float32(float(u1) * np.sqrt((-2.0 * np.log(float(s))) / float(s)))
        recv.append(float(np.float32(x[i] + noise * np.float32(0.7))))
    assert np.array_equal(np.array(host_out["tx_recv"], np.float32), np.array(recv, np.float32))


def test_generator_matrix_pointer(host_out, nr):
    _, _, gen = nr.generator_shifts(1, 384)
    assert host_out["gen_first"] == gen[:3].tolist()
