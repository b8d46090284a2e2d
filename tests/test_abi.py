"""The C-ABI library loads on the CPU and exports every symbol declared in include/*.h."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for f in os.listdir(os.path.join(ROOT, "include")):
        if not f.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", f)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b(ldpc_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_exports_every_declared_symbol(nr):
    names = _declared()
    assert len(names) >= 15
    lib = ctypes.CDLL(nr.LIB_PATH)
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_status_strings(nr):
    assert nr.lib.ldpc_strerror(0) == b"ok"
    assert nr.lib.ldpc_strerror(-1) == b"invalid argument"


def test_host_only_context_refuses_device_work(nr):
    code = nr.Code(1, 384, 0, -1)
    assert code.n_vn == 26112 and code.K == 8448 and code.N == 25344 and code.hd_bytes == 3264
    try:
        code.decode(1, 1, 1, 1)
    except nr.LdpcError as e:
        assert e.status == -4
    else:
        raise AssertionError("decode on a host-only context must fail")


def test_decode_entry_points_refuse_zero_iterations(nr):
    """Every decode entry point takes n_iter >= 1; the argument check runs before any device work."""
    code = nr.Code(1, 384, 0, -1)
    calls = [code.decode, code.decode_half, code.decode_char, code.decode_learned]
    for call in calls:
        for n_iter in (0, -1):
            try:
                call(1, 1, n_iter, 1)
            except nr.LdpcError as e:
                assert e.status == -1, (call.__name__, n_iter, e.status)
            else:
                raise AssertionError(f"{call.__name__} accepted n_iter={n_iter}")


def test_device_encoder_argument_checks(nr):
    """ldpc_encode_async: NULL buffers are refused, a host-only context cannot run it, and an
    empty batch is a no-op; all decided before any device work."""
    code = nr.Code(2, 52, 0, -1)
    assert nr.lib.ldpc_encode_async(code.handle, None, None, 1, ctypes.c_void_p(1)) == -1
    assert nr.lib.ldpc_encode_async(code.handle, None, ctypes.c_void_p(1), -1, ctypes.c_void_p(1)) == -1
    assert nr.lib.ldpc_encode_async(code.handle, None, ctypes.c_void_p(1), 0, ctypes.c_void_p(1)) == 0
    assert nr.lib.ldpc_encode_async(code.handle, None, ctypes.c_void_p(1), 1, ctypes.c_void_p(1)) == -4


def test_cpp_host_api_exported(nr):
    """Every function of include/nr_gpu_ldpc.hpp (the reference-named C++ API) is in the library."""
    import shutil
    import subprocess
    nm = shutil.which("nm")
    if nm is None:
        import pytest
        pytest.skip("nm not available")
    out = subprocess.run([nm, "-DC", "--defined-only", nr.LIB_PATH], capture_output=True, text=True, check=True).stdout
    want = ["setup_ldpc(int, float, int)", "setup_ldpc_qc12x24(int, int)",
            "ldpc_encoder_orig(int*, int*, short, short, unsigned char)", "choose_generator_matrix(short, short)",
            "compute_noise_stddev(float, float, int)", "transmit(int*, int*, float*, unsigned long, float*, int, int, float, int)",
            "calculate_llr(float*, float*, int, float, int)", "error_check(float*, float*, int)",
            "cuda_error_check(int*, char*, int, int, int)", "quantize(double, double, unsigned char)",
            "unpack_char(char*, char*, int)", "host_mem<float> setup_host<float>(params const*, unsigned int, unsigned int)",
            "dev_mem<float> setup_dev_memory<float>(params const*, ldpc_params const*, host_mem<float> const*)",
            "setup_gpu_grid(ldpc_params const*, int, int)",
            "void generate_new_codewords<float>(params const*, host_mem<float>*, ldpc_params const*, int, float)",
            "put_h_compact_and_params_in_constant_memory(ldpc_params const*, params const*, cuda_grid const*)",
            "void perform_gpu_ldpc_decoding<float>(dev_mem<float>*, host_mem<float>*, cuda_grid const*, int, int)",
            "nr_gpu_ldpc_context()", "ldpc_gpu_decoder",
            "host_mem<char> setup_host<char>(params const*, unsigned int, unsigned int)",
            "void perform_gpu_ldpc_decoding<char>(dev_mem<char>*, host_mem<char>*, cuda_grid const*, int, int)",
            "void generate_new_codewords<char>(params const*, host_mem<char>*, ldpc_params const*, int, float)",
            "put_params_in_constant_mem(unsigned int, unsigned int, unsigned int, unsigned int, unsigned int, unsigned int)",
            "nr_gpu_ldpc_fill_compact(h_element*, int, int, h_element*, int, int, ldpc_params const*)",
            "nr_gpu_ldpc_load_compact(h_element const*, int, int, h_element const*, int, int, ldpc_params const*)"]
    missing = [w for w in want if w not in out]
    assert not missing, missing


def test_launch_layout_calls_on_a_host_only_context(nr):
    """Layout calls check their arguments on the CPU: the dup layout's counts are validated
    against its workgroup and LDS limits, the slim layout and the autotuner need a device, and
    the reported shape of a host-only context is the dup layout."""
    code = nr.Code(2, 13, 0, -1)
    cmax = 0
    for c in range(512 // 13, 0, -1):  # the LDS limit is the device's, or a fallback without one
        try:
            code.set_launch_layout(nr.LAYOUT_DUP, c)
            cmax = c
            break
        except nr.LdpcError:
            pass
    assert cmax >= 1
    assert code.launch_layout(100) == (nr.LAYOUT_DUP, cmax, (100 + cmax - 1) // cmax)
    for layout, cpb, status in ((nr.LAYOUT_DUP, cmax + 1, -1), (nr.LAYOUT_SLIM, 0, -4), (7, 0, -1),
                                (nr.LAYOUT_AUTO, 3, -1)):
        try:
            code.set_launch_layout(layout, cpb)
        except nr.LdpcError as e:
            assert e.status == status, (layout, cpb, e.status)
        else:
            raise AssertionError(f"layout {layout} with {cpb} codewords per workgroup was accepted")
    code.set_launch_layout(nr.LAYOUT_AUTO, 0)
    try:
        code.autotune(n_iter=4)
    except nr.LdpcError as e:
        assert e.status == -4
    else:
        raise AssertionError("autotune on a host-only context must fail")


def test_layout_model_calls_on_a_host_only_context(nr):
    """The launch cost model can be read and installed without a device (an autotune result kept
    by the caller and restored in a later process): set/get round trip, the launch shape follows
    the installed dup count, invalid models are refused and leave the model as it was, and the
    slim layout can be dropped from the model and installed again."""
    code = nr.Code(2, 13, 0, -1)
    assert code.layout_model(nr.LAYOUT_DUP)["cpb"] >= 1
    code.set_layout_model(nr.LAYOUT_DUP, 2, 7, 123.5)
    assert code.layout_model(nr.LAYOUT_DUP) == {"cpb": 2, "cw_per_cu": 7, "round_us": 123.5}
    assert code.launch_layout(100) == (nr.LAYOUT_DUP, 2, 50)
    bad = [(nr.LAYOUT_DUP, 0, 1, 1.0), (nr.LAYOUT_DUP, 10 ** 6, 1, 1.0), (nr.LAYOUT_DUP, 1, 0, 1.0),
           (nr.LAYOUT_DUP, 1, 1, 0.0), (nr.LAYOUT_DUP, 1, 1, -3.0), (nr.LAYOUT_DUP, 1, 1, float("nan")),
           (nr.LAYOUT_DUP, 1, 1, float("inf")), (nr.LAYOUT_AUTO, 1, 1, 1.0), (7, 1, 1, 1.0),
           (nr.LAYOUT_SLIM, -1, 1, 1.0), (nr.LAYOUT_SLIM, 10 ** 6, 1, 1.0)]
    for args in bad:
        try:
            code.set_layout_model(*args)
        except nr.LdpcError as e:
            assert e.status == -1, (args, e.status)
        else:
            raise AssertionError(f"model {args} was accepted")
    assert code.layout_model(nr.LAYOUT_DUP) == {"cpb": 2, "cw_per_cu": 7, "round_us": 123.5}
    code.set_layout_model(nr.LAYOUT_SLIM, 0)
    assert code.layout_model(nr.LAYOUT_SLIM) == {"cpb": 0, "cw_per_cu": 0, "round_us": 0.0}
    code.set_layout_model(nr.LAYOUT_SLIM, 1, 3, 50.0)
    assert code.layout_model(nr.LAYOUT_SLIM) == {"cpb": 1, "cw_per_cu": 3, "round_us": 50.0}
    for layout in (0, 3):
        try:
            code.layout_model(layout)
        except nr.LdpcError as e:
            assert e.status == -1
        else:
            raise AssertionError(f"layout {layout} was accepted")
