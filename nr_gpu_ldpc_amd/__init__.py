"""MI355X-native QC-LDPC decoder: Python bindings of the C ABI in include/ldpc_api.h.

The library ``nr_gpu_ldpc_amd/lib/libnrldpc.so`` is built by ``nr_gpu_ldpc_amd/build.py``
(hipcc, gfx950).  Importing this package fails loudly if it has not been built: there is no
CPU or PyTorch fallback for any decode path.

Device buffers are passed as integer device addresses (for example ``tensor.data_ptr()`` of a
torch tensor on ``cuda:0``) and streams as integer hipStream_t handles (0 = default stream).
"""
import ctypes
import os

__all__ = ["lib", "Code", "Mixed", "TransportBlock", "crc", "LdpcError", "BG1", "BG2", "QC12X24", "EARLY_EXIT", "FALLBACK_PLAIN",
           "LEARNED", "MIXED_CW_DTYPE", "setup_params", "base_matrix", "noise_stddev", "generator_shifts",
           "device_count"]

HERE = os.path.dirname(os.path.abspath(__file__))
# NRLDPC_LIB selects another build of the same library (used to compare build variants).
LIB_PATH = os.environ.get("NRLDPC_LIB") or os.path.join(HERE, "lib", "libnrldpc.so")

BG1, BG2, QC12X24 = 1, 2, 3
EARLY_EXIT = 1
FALLBACK_PLAIN = 2
LEARNED = 4
NOALLOC_SCRATCH_BYTES = 32 << 20  # LDPC_NOALLOC_SCRATCH_BYTES of ldpc_api.h
LLR_F32, LLR_I8, LLR_F16 = 0, 1, 2  # LDPC_LLR_* element types of mixed-batch LLRs
CRC16, CRC24A, CRC24B = 0, 1, 2
LAYOUT_AUTO, LAYOUT_DUP, LAYOUT_SLIM = 0, 1, 2  # LDPC_LAYOUT_* launch layouts

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"{LIB_PATH} is missing: build the HIP library first (python nr_gpu_ldpc_amd/build.py)")


def _single_hip_runtime():
    """Load the HIP runtime that PyTorch would load, before libnrldpc.so resolves its own.

    PyTorch-ROCm ships a private libamdhip64 with the same soname as /opt/rocm's.  Whichever
    library loads first decides; loading ours first would leave two HIP runtimes (two device
    enumerations) in one process once torch is imported.  Without torch this is a no-op and
    the system runtime is used.
    """
    import importlib.util
    try:
        spec = importlib.util.find_spec("torch")
    except (ImportError, ValueError):
        spec = None
    if spec is None or not spec.origin:
        return
    hip = os.path.join(os.path.dirname(spec.origin), "lib", "libamdhip64.so")
    if os.path.exists(hip):
        ctypes.CDLL(hip, mode=ctypes.RTLD_GLOBAL)


_single_hip_runtime()
lib = ctypes.CDLL(LIB_PATH)

_c_int_p = ctypes.POINTER(ctypes.c_int)


class CodeInfo(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in
                ("kind", "Z", "ils", "rows", "cols", "kb", "punct", "n_vn", "K", "N", "hd_bytes",
                 "nnz", "cw_per_block", "nnz_full")]


def _sig(name, res, *args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = list(args)
    return f


_P = ctypes.c_void_p
_sig("ldpc_strerror", ctypes.c_char_p, ctypes.c_int)
_sig("ldpc_last_error", ctypes.c_char_p)
_sig("ldpc_device_count", ctypes.c_int, _c_int_p)
_sig("ldpc_ctx_create", ctypes.c_int, ctypes.POINTER(_P), ctypes.c_int, ctypes.c_int, ctypes.c_int,
     ctypes.c_int)
_sig("ldpc_ctx_destroy", ctypes.c_int, _P)
_sig("ldpc_ctx_info", ctypes.c_int, _P, ctypes.POINTER(CodeInfo))
_sig("ldpc_decode_async", ctypes.c_int, _P, _P, _P, ctypes.c_int, ctypes.c_int, ctypes.c_uint, _P, _P, _P)
_sig("ldpc_set_launch", ctypes.c_int, _P, ctypes.c_int)
_sig("ldpc_launch_shape", ctypes.c_int, _P, ctypes.c_int, _P, _P)
_sig("ldpc_set_launch_layout", ctypes.c_int, _P, ctypes.c_int, ctypes.c_int)
_sig("ldpc_launch_layout", ctypes.c_int, _P, ctypes.c_int, _P, _P, _P)
_sig("ldpc_autotune", ctypes.c_int, _P, ctypes.c_int, ctypes.c_int, _P, _P, _P, _P, _P)
_sig("ldpc_get_layout_model", ctypes.c_int, _P, ctypes.c_int, _P, _P, _P)
_sig("ldpc_set_layout_model", ctypes.c_int, _P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float)
_sig("ldpc_syndrome_async", ctypes.c_int, _P, _P, _P, ctypes.c_int, _P)
_sig("ldpc_hd_to_stream_async", ctypes.c_int, _P, _P, _P, ctypes.c_int, _P)
_sig("ldpc_decode_half_async", ctypes.c_int, _P, _P, _P, ctypes.c_int, ctypes.c_int, ctypes.c_uint, _P, _P, _P)
_sig("ldpc_decode_char_async", ctypes.c_int, _P, _P, _P, ctypes.c_int, ctypes.c_int, ctypes.c_uint, _P, _P, _P)
_sig("ldpc_llr_to_half_async", ctypes.c_int, _P, _P, ctypes.c_size_t, _P)
_sig("ldpc_base_matrix", ctypes.c_int, ctypes.c_int, ctypes.c_int, _c_int_p, ctypes.c_int)
_sig("ldpc_setup_params", ctypes.c_int, ctypes.c_int, ctypes.c_float, _c_int_p)
_sig("ldpc_encode_host", ctypes.c_int, _P, _P, ctypes.c_int, _P)
_sig("ldpc_encode_async", ctypes.c_int, _P, _P, _P, ctypes.c_int, _P)
_sig("ldpc_generator_shifts", ctypes.c_int, ctypes.c_int, ctypes.c_int, _P, _P, _P, ctypes.c_int)
_sig("ldpc_noise_stddev", ctypes.c_float, ctypes.c_float, ctypes.c_float)
_sig("ldpc_channel_async", ctypes.c_int, _P, _P, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int, _P,
     ctypes.c_float, _P, _P)
_sig("ldpc_count_errors_async", ctypes.c_int, _P, _P, _P, _P, ctypes.c_int, _P)
_sig("ldpc_noms_set_weights", ctypes.c_int, _P, ctypes.c_int, _P)
_sig("ldpc_noms_get_weights", ctypes.c_int, _P, _c_int_p, _P)
_sig("ldpc_decode_learned_async", ctypes.c_int, _P, _P, _P, ctypes.c_int, ctypes.c_int, ctypes.c_uint, _P, _P, _P)
_sig("ldpc_decode_learned_noalloc_async", ctypes.c_int, _P, _P, _P, ctypes.c_int, ctypes.c_int, ctypes.c_uint, _P, _P, _P)
_sig("ldpc_noms_workspace_bytes", ctypes.c_size_t, _P, ctypes.c_int, ctypes.c_int)
_sig("ldpc_noms_train_grad_weighted_async", ctypes.c_int, _P, _P, _P, _P, _P, ctypes.c_int, ctypes.c_int, _P,
     _P, _P, ctypes.c_size_t)
_sig("ldpc_noms_train_grad_async", ctypes.c_int, _P, _P, _P, _P, ctypes.c_int, ctypes.c_int, _P, _P, _P,
     ctypes.c_size_t)
_sig("ldpc_noms_adam_async", ctypes.c_int, _P, _P, _P, _P, _P, ctypes.c_int, ctypes.c_float)
_sig("ldpc_mixed_create", ctypes.c_int, ctypes.POINTER(_P), ctypes.c_int, ctypes.c_int)
_sig("ldpc_mixed_destroy", ctypes.c_int, _P)
_sig("ldpc_mixed_plan", ctypes.c_int, _P, _P, ctypes.c_int)
_sig("ldpc_mixed_plan_info", ctypes.c_int, _P, _c_int_p, _c_int_p, _c_int_p, _c_int_p)
_sig("ldpc_mixed_tiles", ctypes.c_int, _P, _P, ctypes.c_int, _P)
_sig("ldpc_mixed_decode_async", ctypes.c_int, _P, _P, _P, ctypes.c_int, ctypes.c_uint, _P, _P, _P)
_sig("ldpc_mixed_capture", ctypes.c_int, _P, _P, _P, ctypes.c_int, ctypes.c_uint, _P, _P, _P, _P, ctypes.c_size_t,
     _P, ctypes.c_size_t)
_sig("ldpc_mixed_replay", ctypes.c_int, _P, _P)
_sig("ldpc_mixed_set_input", ctypes.c_int, _P, ctypes.c_int, ctypes.c_float)
_sig("ldpc_mixed_set_grid", ctypes.c_int, _P, ctypes.c_int)
_sig("ldpc_mixed_set_weights", ctypes.c_int, _P, ctypes.c_int, ctypes.c_int, ctypes.c_int, _P)
_sig("ldpc_mixed_set_layout", ctypes.c_int, _P, ctypes.c_int)
_sig("ldpc_mixed_get_layout", ctypes.c_int, _P, _c_int_p)
_sig("ldpc_mixed_autotune_layout", ctypes.c_int, _P, _P, ctypes.c_int, _P, ctypes.c_int, ctypes.c_uint, _P,
     ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float), _c_int_p)
_sig("ldpc_mixed_set_cpb", ctypes.c_int, _P, ctypes.c_int, ctypes.c_int, ctypes.c_int)
_sig("ldpc_mixed_get_cpb", ctypes.c_int, _P, ctypes.c_int, ctypes.c_int, _c_int_p)
_sig("ldpc_mixed_autotune", ctypes.c_int, _P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _P, _P,
     ctypes.c_int, _c_int_p, _c_int_p)
_sig("ldpc_mixed_save_tuning", ctypes.c_int, _P, ctypes.c_char_p)
_sig("ldpc_mixed_load_tuning", ctypes.c_int, _P, ctypes.c_char_p)
_sig("ldpc_crc", ctypes.c_int, ctypes.c_int, _P, ctypes.c_longlong, ctypes.POINTER(ctypes.c_uint32))
_sig("ldpc_tb_create", ctypes.c_int, ctypes.POINTER(_P), ctypes.c_int, ctypes.c_float, ctypes.c_int, ctypes.c_int,
     ctypes.c_int, ctypes.c_int, ctypes.c_int)
_sig("ldpc_tb_destroy", ctypes.c_int, _P)
_sig("ldpc_tb_get_info", ctypes.c_int, _P, _P, _c_int_p, ctypes.c_int)
_sig("ldpc_tb_encode_host", ctypes.c_int, _P, _P, _P, _P)
_sig("ldpc_tb_derate_async", ctypes.c_int, _P, _P, _P, _P)
_sig("ldpc_tb_check_async", ctypes.c_int, _P, _P, _P, _P, _P)


class LdpcError(RuntimeError):
    def __init__(self, status, what):
        msg = lib.ldpc_strerror(status).decode()
        detail = lib.ldpc_last_error().decode()
        super().__init__(f"{what}: {msg}" + (f" ({detail})" if detail else ""))
        self.status = status


def _check(status, what):
    if status != 0:
        raise LdpcError(status, what)


def llr_to_half(llr_ptr, n, out_ptr, stream=0):
    """Device float32 -> binary16 (round to nearest even), n values."""
    _check(lib.ldpc_llr_to_half_async(stream, llr_ptr, n, out_ptr), "ldpc_llr_to_half_async")


def device_count():
    n = ctypes.c_int()
    _check(lib.ldpc_device_count(ctypes.byref(n)), "ldpc_device_count")
    return n.value


def setup_params(A, rate):
    """setup_ldpc (src/nr_ldpc.cpp:31-230): dict of the derived code parameters."""
    out = (ctypes.c_int * 10)()
    _check(lib.ldpc_setup_params(int(A), float(rate), out), "ldpc_setup_params")
    keys = ("bg", "Z", "K", "N", "H_rows", "H_cols", "ils", "N_before_puncture", "n_nz_in_row",
            "n_nz_in_col")
    return dict(zip(keys, list(out)))


def base_matrix(kind, ils):
    import numpy as np
    rows, cols = {BG1: (46, 68), BG2: (42, 52), QC12X24: (12, 24)}[kind]
    out = np.empty((rows, cols), dtype=np.int32)
    _check(lib.ldpc_base_matrix(kind, ils, out.ctypes.data_as(_c_int_p), out.size), "ldpc_base_matrix")
    return out


def noise_stddev(ebn0_db, rate):
    return lib.ldpc_noise_stddev(float(ebn0_db), float(rate))


def generator_shifts(kind, Z):
    """Regenerated OAI-layout generator tables: (no_shift, pointer_shift, gen_shift) int16 arrays."""
    import numpy as np
    n = lib.ldpc_generator_shifts(kind, Z, None, None, None, 0)
    if n < 0:
        _check(n, "ldpc_generator_shifts")
    kb = 22 if kind == BG1 else 10
    rows = 46 if kind == BG1 else 42
    no = np.zeros(rows * kb, dtype=np.int16)
    ptr = np.zeros(rows * kb, dtype=np.int16)
    gen = np.zeros(n, dtype=np.int16)
    lib.ldpc_generator_shifts(kind, Z, no.ctypes.data, ptr.ctypes.data, gen.ctypes.data, n)
    return no, ptr, gen


class Code:
    """One decoder context (code kind, lifting size Z, kept parity rows) on one device."""

    def __init__(self, kind, Z, m_rows=0, device=0):
        h = _P()
        _check(lib.ldpc_ctx_create(ctypes.byref(h), kind, Z, m_rows, device), "ldpc_ctx_create")
        self._h = h
        info = CodeInfo()
        _check(lib.ldpc_ctx_info(h, ctypes.byref(info)), "ldpc_ctx_info")
        self.info = {f: getattr(info, f) for f, _ in CodeInfo._fields_}
        for k, v in self.info.items():
            setattr(self, k, v)

    def close(self):
        if getattr(self, "_h", None):
            lib.ldpc_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def set_launch(self, cw_per_block):
        _check(lib.ldpc_set_launch(self._h, int(cw_per_block)), "ldpc_set_launch")
        self.cw_per_block = cw_per_block or self.info["cw_per_block"]

    def launch_shape(self, n_cw):
        """(codewords per workgroup, workgroups) of a plain or learned decode of n_cw codewords."""
        cpb, wg = ctypes.c_int(0), ctypes.c_int(0)
        _check(lib.ldpc_launch_shape(self._h, int(n_cw), ctypes.byref(cpb), ctypes.byref(wg)), "ldpc_launch_shape")
        return cpb.value, wg.value

    def set_launch_layout(self, layout, cw_per_block=0):
        """LAYOUT_AUTO (per batch), LAYOUT_DUP or LAYOUT_SLIM with codewords per workgroup (0: tuned)."""
        _check(lib.ldpc_set_launch_layout(self._h, int(layout), int(cw_per_block)), "ldpc_set_launch_layout")

    def launch_layout(self, n_cw):
        """(layout, codewords per workgroup, workgroups) of a plain or learned decode of n_cw codewords."""
        lay, cpb, wg = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(0)
        _check(lib.ldpc_launch_layout(self._h, int(n_cw), ctypes.byref(lay), ctypes.byref(cpb), ctypes.byref(wg)),
               "ldpc_launch_layout")
        return lay.value, cpb.value, wg.value

    def autotune(self, n_iter=10, max_candidates=64):
        """Time every launch shape and install the fastest of each layout; returns the candidates as
        dicts (layout, cpb, cw_per_cu, round_us)."""
        n = int(max_candidates)
        lay, cpb, per = (ctypes.c_int * n)(), (ctypes.c_int * n)(), (ctypes.c_int * n)()
        rus, cnt = (ctypes.c_float * n)(), ctypes.c_int(0)
        _check(lib.ldpc_autotune(self._h, int(n_iter), n, lay, cpb, per, rus, ctypes.byref(cnt)), "ldpc_autotune")
        return [{"layout": lay[k], "cpb": cpb[k], "cw_per_cu": per[k], "round_us": rus[k]}
                for k in range(min(n, cnt.value))]

    def layout_model(self, layout):
        """The launch cost model of LAYOUT_DUP or LAYOUT_SLIM as a dict (cpb, cw_per_cu, round_us);
        cpb 0: the layout is not in the model."""
        cpb, per, rus = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_float(0.0)
        _check(lib.ldpc_get_layout_model(self._h, int(layout), ctypes.byref(cpb), ctypes.byref(per),
                                         ctypes.byref(rus)), "ldpc_get_layout_model")
        return {"cpb": cpb.value, "cw_per_cu": per.value, "round_us": rus.value}

    def set_layout_model(self, layout, cpb, cw_per_cu=0, round_us=0.0):
        """Install a launch cost model measured earlier (an autotune cache); cpb 0 drops the slim
        layout from the model."""
        _check(lib.ldpc_set_layout_model(self._h, int(layout), int(cpb), int(cw_per_cu), float(round_us)),
               "ldpc_set_layout_model")

    def decode(self, llr_ptr, n_cw, n_iter, hd_ptr, app_ptr=0, iters_ptr=0, flags=0, stream=0):
        _check(lib.ldpc_decode_async(self._h, stream, llr_ptr, n_cw, n_iter, flags, hd_ptr,
                                     app_ptr or None, iters_ptr or None), "ldpc_decode_async")

    def decode_half(self, llr_half_ptr, n_cw, n_iter, hd_ptr, app_ptr=0, iters_ptr=0, flags=0, stream=0):
        """The reference's HALF_PRC schedule; llr_half_ptr: device binary16 [n_cw][n_vn]."""
        _check(lib.ldpc_decode_half_async(self._h, stream, llr_half_ptr, n_cw, n_iter, flags, hd_ptr,
                                          app_ptr or None, iters_ptr or None), "ldpc_decode_half_async")

    def decode_char(self, llr_i8_ptr, n_cw, n_iter, hd_ptr, app_ptr=0, iters_ptr=0, flags=0, stream=0):
        """The reference's CHAR_PRC schedule; llr_i8_ptr: device int8 [n_cw][n_vn]."""
        _check(lib.ldpc_decode_char_async(self._h, stream, llr_i8_ptr, n_cw, n_iter, flags, hd_ptr,
                                          app_ptr or None, iters_ptr or None), "ldpc_decode_char_async")

    def syndrome(self, hd_ptr, n_cw, weight_ptr, stream=0):
        _check(lib.ldpc_syndrome_async(self._h, stream, hd_ptr, n_cw, weight_ptr), "ldpc_syndrome_async")

    def hd_to_stream(self, hd_ptr, n_cw, out_ptr, stream=0):
        _check(lib.ldpc_hd_to_stream_async(self._h, stream, hd_ptr, n_cw, out_ptr), "ldpc_hd_to_stream_async")

    def channel(self, seed, stream_base, n_cw, llr_ptr, sigma=None, sigma_ptr=0, cw_bits_ptr=0, stream=0):
        _check(lib.ldpc_channel_async(self._h, stream, seed, stream_base, n_cw, sigma_ptr or None,
                                      float(sigma or 0.0), cw_bits_ptr or None, llr_ptr),
               "ldpc_channel_async")

    def count_errors(self, hd_ptr, cw_bits_ptr, n_cw, counts_ptr, stream=0):
        _check(lib.ldpc_count_errors_async(self._h, stream, hd_ptr, cw_bits_ptr, n_cw, counts_ptr),
               "ldpc_count_errors_async")

    # ---- learned offset/normalized min-sum
    def set_weights(self, weights):
        """weights: float32 [n_iter, nnz_full, 2] (alpha, beta)."""
        import numpy as np
        w = np.ascontiguousarray(weights, dtype=np.float32)
        assert w.ndim == 3 and w.shape[1:] == (self.nnz_full, 2)
        _check(lib.ldpc_noms_set_weights(self._h, w.shape[0], w.ctypes.data), "ldpc_noms_set_weights")

    def get_weights(self):
        import numpy as np
        t = ctypes.c_int()
        _check(lib.ldpc_noms_get_weights(self._h, ctypes.byref(t), None), "ldpc_noms_get_weights")
        w = np.empty((t.value, self.nnz_full, 2), dtype=np.float32)
        _check(lib.ldpc_noms_get_weights(self._h, ctypes.byref(t), w.ctypes.data), "ldpc_noms_get_weights")
        return w

    def plain_weights(self, n_iter):
        import numpy as np
        w = np.zeros((n_iter, self.nnz_full, 2), dtype=np.float32)
        w[..., 0] = 0.75
        return w

    def decode_learned(self, llr_ptr, n_cw, n_iter, hd_ptr, app_ptr=0, iters_ptr=0, flags=0, stream=0):
        _check(lib.ldpc_decode_learned_async(self._h, stream, llr_ptr, n_cw, n_iter, flags, hd_ptr,
                                             app_ptr or None, iters_ptr or None), "ldpc_decode_learned_async")

    def decode_learned_noalloc(self, llr_ptr, n_cw, n_iter, hd_ptr, app_ptr=0, iters_ptr=0, flags=0, stream=0):
        """decode_learned with the fallback's temporaries in the library's fixed per-GPU array
        (nothing is allocated; at most NOALLOC_SCRATCH_BYTES of temporaries)."""
        _check(lib.ldpc_decode_learned_noalloc_async(self._h, stream, llr_ptr, n_cw, n_iter, flags, hd_ptr,
                                                      app_ptr or None, iters_ptr or None),
               "ldpc_decode_learned_noalloc_async")

    def workspace_bytes(self, n_cw, n_iter):
        return lib.ldpc_noms_workspace_bytes(self._h, n_cw, n_iter)

    def train_grad(self, llr_ptr, bits_ptr, n_cw, n_iter, grad_ptr, loss_ptr, ws_ptr, ws_bytes, stream=0,
                   weight_ptr=0):
        """weight_ptr: device float [n_cw] loss weight per codeword (0: all 1)."""
        if weight_ptr:
            _check(lib.ldpc_noms_train_grad_weighted_async(self._h, stream, llr_ptr, bits_ptr, weight_ptr, n_cw,
                                                           n_iter, grad_ptr, loss_ptr, ws_ptr, ws_bytes),
                   "ldpc_noms_train_grad_weighted_async")
            return
        _check(lib.ldpc_noms_train_grad_async(self._h, stream, llr_ptr, bits_ptr, n_cw, n_iter, grad_ptr,
                                              loss_ptr, ws_ptr, ws_bytes), "ldpc_noms_train_grad_async")

    def adam(self, grad_ptr, m_ptr, v_ptr, step, lr, stream=0):
        _check(lib.ldpc_noms_adam_async(self._h, stream, grad_ptr, m_ptr, v_ptr, step, lr), "ldpc_noms_adam_async")

    def encode(self, info_ptr, n_cw, cw_ptr, stream=0):
        """Device encoder: info bits [n_cw][K] (uint8, device) -> codeword bits [n_cw][n_vn]."""
        _check(lib.ldpc_encode_async(self._h, stream, info_ptr, n_cw, cw_ptr), "ldpc_encode_async")

    def encode_host(self, info_bits):
        """info_bits: uint8 array [n_cw, K] -> codeword bits uint8 [n_cw, n_vn] (host)."""
        import numpy as np
        info_bits = np.ascontiguousarray(info_bits, dtype=np.uint8).reshape(-1, self.K)
        out = np.empty((info_bits.shape[0], self.n_vn), dtype=np.uint8)
        _check(lib.ldpc_encode_host(self._h, info_bits.ctypes.data, info_bits.shape[0], out.ctypes.data),
               "ldpc_encode_host")
        return out


def _mixed_cw_dtype():
    import numpy as np
    return np.dtype([("kind", "<i4"), ("Z", "<i4"), ("m_rows", "<i4"), ("reserved", "<i4"),
                     ("llr_offset", "<i8"), ("hd_offset", "<i8")])


try:
    MIXED_CW_DTYPE = _mixed_cw_dtype()  # layout of ldpc_mixed_cw
except ImportError:  # numpy absent: descriptors cannot be built from Python
    MIXED_CW_DTYPE = None


class Mixed:
    """Persistent mixed-batch decoder (base graph, Z and parity rows per codeword).

    Descriptors are a numpy array of MIXED_CW_DTYPE (kind, Z, m_rows, 0, llr_offset, hd_offset).
    device < 0 makes a host-only planner (plan and tiles work, decoding does not)."""

    def __init__(self, max_cw, device=0):
        h = _P()
        _check(lib.ldpc_mixed_create(ctypes.byref(h), device, int(max_cw)), "ldpc_mixed_create")
        self._h = h
        self.max_cw = max_cw

    def close(self):
        if getattr(self, "_h", None):
            lib.ldpc_mixed_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def plan(self, cws):
        import numpy as np
        self._cws = np.ascontiguousarray(cws, dtype=MIXED_CW_DTYPE)
        _check(lib.ldpc_mixed_plan(self._h, self._cws.ctypes.data, len(self._cws)), "ldpc_mixed_plan")

    def plan_info(self):
        v = [ctypes.c_int() for _ in range(4)]
        _check(lib.ldpc_mixed_plan_info(self._h, *[ctypes.byref(x) for x in v]), "ldpc_mixed_plan_info")
        return dict(zip(("n_tiles", "threads", "grid", "lds_bytes"), (x.value for x in v)))

    def tiles(self):
        """(tiles [n_tiles, 6] = kind, Z, m_rows, cpb, first, count; order [n_cw])."""
        import numpy as np
        n = self.plan_info()["n_tiles"]
        out = np.zeros((max(n, 1), 6), dtype=np.int32)
        order = np.zeros(max(len(self._cws), 1), dtype=np.int32)
        r = lib.ldpc_mixed_tiles(self._h, out.ctypes.data, n, order.ctypes.data)
        if r < 0:
            _check(r, "ldpc_mixed_tiles")
        return out[:n], order[:len(self._cws)]

    def decode(self, llr_ptr, n_iter, hd_ptr, app_ptr=0, iters_ptr=0, flags=0, stream=0):
        _check(lib.ldpc_mixed_decode_async(self._h, stream, llr_ptr, n_iter, flags, hd_ptr, app_ptr or None,
                                           iters_ptr or None), "ldpc_mixed_decode_async")

    def capture(self, stream, llr_ptr, n_iter, hd_ptr, app_ptr=0, iters_ptr=0, flags=0, llr_host_ptr=0,
                llr_floats=0, hd_host_ptr=0, hd_bytes=0):
        _check(lib.ldpc_mixed_capture(self._h, stream, llr_ptr, n_iter, flags, hd_ptr, app_ptr or None,
                                      iters_ptr or None, llr_host_ptr or None, llr_floats, hd_host_ptr or None,
                                      hd_bytes), "ldpc_mixed_capture")

    def replay(self, stream):
        _check(lib.ldpc_mixed_replay(self._h, stream), "ldpc_mixed_replay")

    def set_input(self, fmt, scale=1.0):
        """Element type of the LLR buffers of later decodes/captures: LLR_F32, LLR_I8 (q * scale)
        or LLR_F16 (binary16 bit patterns)."""
        _check(lib.ldpc_mixed_set_input(self._h, fmt, scale), "ldpc_mixed_set_input")

    def set_grid(self, workgroups):
        """Persistent-grid workgroups of later launches/captures (0: the whole device)."""
        _check(lib.ldpc_mixed_set_grid(self._h, workgroups), "ldpc_mixed_set_grid")

    def set_weights(self, kind, Z, weights):
        import numpy as np
        w = np.ascontiguousarray(weights, dtype=np.float32)
        _check(lib.ldpc_mixed_set_weights(self._h, kind, Z, w.shape[0], w.ctypes.data), "ldpc_mixed_set_weights")

    def set_layout(self, layout):
        """LAYOUT_DUP, LAYOUT_SLIM or LAYOUT_AUTO for the tiles of later plans (discards the plan and
        any captured graph)."""
        _check(lib.ldpc_mixed_set_layout(self._h, int(layout)), "ldpc_mixed_set_layout")

    def autotune_layout(self, cws, llr_ptr, n_iter, hd_ptr, flags=0):
        """Times both layouts on a representative batch (descriptors + device LLR / HD buffers),
        keeps the faster; returns (best layout, {layout: microseconds per batch})."""
        import numpy as np
        self._cws = np.ascontiguousarray(cws, dtype=MIXED_CW_DTYPE)
        ud, us, best = ctypes.c_float(), ctypes.c_float(), ctypes.c_int()
        _check(lib.ldpc_mixed_autotune_layout(self._h, self._cws.ctypes.data, len(self._cws), llr_ptr, n_iter, flags,
                                              hd_ptr, ctypes.byref(ud), ctypes.byref(us), ctypes.byref(best)),
               "ldpc_mixed_autotune_layout")
        return best.value, {LAYOUT_DUP: ud.value, LAYOUT_SLIM: us.value}

    def get_layout(self):
        v = ctypes.c_int()
        _check(lib.ldpc_mixed_get_layout(self._h, ctypes.byref(v)), "ldpc_mixed_get_layout")
        return v.value

    def set_cpb(self, kind, Z, cpb):
        _check(lib.ldpc_mixed_set_cpb(self._h, kind, Z, cpb), "ldpc_mixed_set_cpb")

    def get_cpb(self, kind, Z):
        c = ctypes.c_int()
        _check(lib.ldpc_mixed_get_cpb(self._h, kind, Z, ctypes.byref(c)), "ldpc_mixed_get_cpb")
        return c.value

    def autotune(self, kind, Z, n_iter, n_cw):
        """Times every candidate codewords-per-workgroup count; returns (best, {cpb: us_per_cw})."""
        import numpy as np
        cand = np.zeros(16, dtype=np.int32)
        us = np.zeros(16, dtype=np.float32)
        n, best = ctypes.c_int(), ctypes.c_int()
        _check(lib.ldpc_mixed_autotune(self._h, kind, Z, n_iter, n_cw, cand.ctypes.data, us.ctypes.data, 16,
                                       ctypes.byref(n), ctypes.byref(best)), "ldpc_mixed_autotune")
        k = min(n.value, 16)
        return best.value, {int(c): float(u) for c, u in zip(cand[:k], us[:k])}

    def save_tuning(self, path):
        _check(lib.ldpc_mixed_save_tuning(self._h, os.fsencode(path)), "ldpc_mixed_save_tuning")

    def load_tuning(self, path):
        _check(lib.ldpc_mixed_load_tuning(self._h, os.fsencode(path)), "ldpc_mixed_load_tuning")


def crc(kind, bits):
    """CRC parity (TS 38.212 5.1) of a 0/1 bit sequence, as an int (p_0 = most significant)."""
    import numpy as np
    b = np.ascontiguousarray(bits, dtype=np.uint8)
    out = ctypes.c_uint32()
    _check(lib.ldpc_crc(kind, b.ctypes.data, b.size, ctypes.byref(out)), "ldpc_crc")
    return out.value


class TbInfo(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in ("A", "L_tb", "B", "bg", "C", "L_cb", "K_prime", "K", "Z", "F", "N",
                                              "n_vn", "G", "Qm", "rv", "k0")]


class TransportBlock:
    """A transport block of A bits, segmented, LDPC-coded and rate-matched to G bits (TS 38.212).
    device < 0: host only (encoding)."""

    def __init__(self, A, rate, G, Qm=2, rv=0, bg=0, device=0):
        import numpy as np
        h = _P()
        _check(lib.ldpc_tb_create(ctypes.byref(h), A, rate, bg, G, Qm, rv, device), "ldpc_tb_create")
        self._h = h
        inf = TbInfo()
        E = np.zeros(4096, dtype=np.int32)
        _check(lib.ldpc_tb_get_info(h, ctypes.byref(inf), E.ctypes.data_as(_c_int_p), E.size), "ldpc_tb_get_info")
        self.info = {f: getattr(inf, f) for f, _ in TbInfo._fields_}
        for k, v in self.info.items():
            setattr(self, k, v)
        self.E = E[:self.C].copy()

    def close(self):
        if getattr(self, "_h", None):
            lib.ldpc_tb_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def encode_host(self, a_bits):
        """payload bits [A] -> (transmitted bits [G], code blocks [C][n_vn])."""
        import numpy as np
        a = np.ascontiguousarray(a_bits, dtype=np.uint8)
        assert a.size == self.A
        f = np.empty(self.G, dtype=np.uint8)
        cw = np.empty((self.C, self.n_vn), dtype=np.uint8)
        _check(lib.ldpc_tb_encode_host(self._h, a.ctypes.data, f.ctypes.data, cw.ctypes.data), "ldpc_tb_encode_host")
        return f, cw

    def derate(self, llr_in_ptr, llr_cb_ptr, stream=0):
        _check(lib.ldpc_tb_derate_async(self._h, stream, llr_in_ptr, llr_cb_ptr), "ldpc_tb_derate_async")

    def check(self, hd_ptr, cb_ok_ptr, tb_bits_ptr=0, stream=0):
        _check(lib.ldpc_tb_check_async(self._h, stream, hd_ptr, cb_ok_ptr, tb_bits_ptr or None), "ldpc_tb_check_async")
