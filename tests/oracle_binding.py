"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so) -- test infrastructure.

The oracle is the checker only: tests compare the product library against it.
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "_build", "liboracle.so")

OR_BG1, OR_BG2, OR_WIFI = 1, 2, 3


class OracleCode(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("Z", ctypes.c_int), ("ils", ctypes.c_int),
                ("rows", ctypes.c_int), ("cols", ctypes.c_int), ("rows_full", ctypes.c_int),
                ("cols_full", ctypes.c_int), ("kb", ctypes.c_int), ("V", ctypes.POINTER(ctypes.c_int))]


class Oracle:
    def __init__(self, lib):
        self.lib = lib
        P = ctypes.c_void_p
        lib.oracle_code_init.argtypes = [ctypes.POINTER(OracleCode), ctypes.c_int, ctypes.c_int, ctypes.c_int]
        lib.oracle_code_free.argtypes = [ctypes.POINTER(OracleCode)]
        lib.oracle_decode.argtypes = [ctypes.POINTER(OracleCode), P, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      P, P, P, ctypes.c_int]
        lib.oracle_decode_noms.argtypes = [ctypes.POINTER(OracleCode), P, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                           P, P, P, P, P, ctypes.c_int]
        lib.oracle_decode_half.argtypes = [ctypes.POINTER(OracleCode), P, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                           P, P, P, ctypes.c_int]
        lib.oracle_decode_char.argtypes = [ctypes.POINTER(OracleCode), P, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                           P, P, P, ctypes.c_int]
        lib.oracle_decode_char.restype = ctypes.c_longlong
        lib.oracle_float_to_half.argtypes = [ctypes.c_float]
        lib.oracle_float_to_half.restype = ctypes.c_uint16
        lib.oracle_half_to_float.argtypes = [ctypes.c_uint16]
        lib.oracle_half_to_float.restype = ctypes.c_float
        lib.oracle_noms_loss_grad.argtypes = [ctypes.POINTER(OracleCode), P, P, ctypes.c_int, ctypes.c_int,
                                              P, P, P, P]
        lib.oracle_noms_loss_grad.restype = ctypes.c_double
        lib.oracle_noms_loss_grad_weighted.argtypes = [ctypes.POINTER(OracleCode), P, P, P, ctypes.c_int,
                                                       ctypes.c_int, P, P, P, P]
        lib.oracle_noms_loss_grad_weighted.restype = ctypes.c_double
        lib.oracle_syndrome.argtypes = [ctypes.POINTER(OracleCode), P]
        lib.oracle_noise_stddev.argtypes = [ctypes.c_float, ctypes.c_float]
        lib.oracle_noise_stddev.restype = ctypes.c_float
        lib.oracle_llr.argtypes = [P, ctypes.c_int, ctypes.c_float, ctypes.c_int, P]
        lib.oracle_gauss.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64]
        lib.oracle_gauss.restype = ctypes.c_float
        lib.oracle_info_bit_word.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64]
        lib.oracle_info_bit_word.restype = ctypes.c_uint32
        lib.oracle_setup_ldpc.argtypes = [ctypes.c_int, ctypes.c_float, P]
        lib.oracle_base_matrix.argtypes = [ctypes.c_int, ctypes.c_int, P]

    def code(self, kind, Z, m_rows=0):
        c = OracleCode()
        if self.lib.oracle_code_init(ctypes.byref(c), kind, Z, m_rows) != 0:
            raise ValueError((kind, Z, m_rows))
        return c

    def shifts(self, c):
        """Dense V mod Z (-1 = zero block) restricted to the rows/columns in use."""
        v = np.ctypeslib.as_array(c.V, shape=(c.rows_full * c.cols_full,)).reshape(c.rows_full, c.cols_full)
        return v[:c.rows, :c.cols].copy()

    def decode(self, c, llr, n_iter, early_exit=False, threads=8):
        llr = np.ascontiguousarray(llr, dtype=np.float32)
        n_cw = llr.shape[0]
        nvn = c.cols * c.Z
        assert llr.shape[1] == nvn
        hd = np.empty((n_cw, nvn), dtype=np.uint8)
        app = np.empty((n_cw, nvn), dtype=np.float32)
        iters = np.empty(n_cw, dtype=np.int32)
        self.lib.oracle_decode(ctypes.byref(c), llr.ctypes.data, n_cw, n_iter, int(early_exit), hd.ctypes.data,
                               app.ctypes.data, iters.ctypes.data, threads)
        return hd, app, iters

    def decode_half(self, c, llr_half, n_iter, early_exit=False, threads=8):
        """llr_half: uint16 (IEEE binary16 bit patterns) [n_cw][cols*Z]."""
        llr_half = np.ascontiguousarray(llr_half, dtype=np.uint16)
        n_cw, nvn = llr_half.shape
        hd = np.empty((n_cw, nvn), dtype=np.uint8)
        app = np.empty((n_cw, nvn), dtype=np.float32)
        iters = np.empty(n_cw, dtype=np.int32)
        self.lib.oracle_decode_half(ctypes.byref(c), llr_half.ctypes.data, n_cw, n_iter, int(early_exit),
                                    hd.ctypes.data, app.ctypes.data, iters.ctypes.data, threads)
        return hd, app, iters

    def decode_char(self, c, llr_i8, n_iter, early_exit=False, threads=8):
        """llr_i8: int8 LLRs [n_cw][cols*Z] (CHAR_PRC).  Returns hd, app, iters and the number of
        out-of-range float -> char conversions (saturated)."""
        llr_i8 = np.ascontiguousarray(llr_i8, dtype=np.int8)
        n_cw, nvn = llr_i8.shape
        hd = np.empty((n_cw, nvn), dtype=np.uint8)
        app = np.empty((n_cw, nvn), dtype=np.float32)
        iters = np.empty(n_cw, dtype=np.int32)
        oor = self.lib.oracle_decode_char(ctypes.byref(c), llr_i8.ctypes.data, n_cw, n_iter, int(early_exit),
                                          hd.ctypes.data, app.ctypes.data, iters.ctypes.data, threads)
        return hd, app, iters, int(oor)

    def decode_noms(self, c, llr, n_iter, alpha, beta, early_exit=False, threads=8, with_iters=False):
        """Learned decode; alpha, beta: [n_iter][nnz of the rows in use] (flattened).  Returns
        hd, app (and iters when with_iters)."""
        llr = np.ascontiguousarray(llr, dtype=np.float32)
        n_cw, nvn = llr.shape
        hd = np.empty((n_cw, nvn), dtype=np.uint8)
        app = np.empty((n_cw, nvn), dtype=np.float32)
        iters = np.empty(n_cw, dtype=np.int32)
        alpha = np.ascontiguousarray(alpha, dtype=np.float32)
        beta = np.ascontiguousarray(beta, dtype=np.float32)
        self.lib.oracle_decode_noms(ctypes.byref(c), llr.ctypes.data, n_cw, n_iter, int(early_exit),
                                    alpha.ctypes.data, beta.ctypes.data, hd.ctypes.data, app.ctypes.data,
                                    iters.ctypes.data, threads)
        return (hd, app, iters) if with_iters else (hd, app)

    def noms_loss_grad(self, c, llr, bits, n_iter, alpha, beta, grad=True, cw_weight=None):
        """Mean BCE of the learned decode and its gradient; cw_weight [n_cw] weights each codeword's
        loss terms (None: the unweighted entry point)."""
        llr = np.ascontiguousarray(llr, dtype=np.float32)
        bits = np.ascontiguousarray(bits, dtype=np.uint8)
        alpha = np.ascontiguousarray(alpha, dtype=np.float64)
        beta = np.ascontiguousarray(beta, dtype=np.float64)
        ga = np.zeros_like(alpha) if grad else None
        gb = np.zeros_like(beta) if grad else None
        gpa = ga.ctypes.data if grad else None
        gpb = gb.ctypes.data if grad else None
        if cw_weight is None:
            loss = self.lib.oracle_noms_loss_grad(ctypes.byref(c), llr.ctypes.data, bits.ctypes.data, llr.shape[0],
                                                  n_iter, alpha.ctypes.data, beta.ctypes.data, gpa, gpb)
        else:
            wt = np.ascontiguousarray(cw_weight, dtype=np.float64)
            loss = self.lib.oracle_noms_loss_grad_weighted(ctypes.byref(c), llr.ctypes.data, bits.ctypes.data,
                                                           wt.ctypes.data, llr.shape[0], n_iter, alpha.ctypes.data,
                                                           beta.ctypes.data, gpa, gpb)
        return loss, ga, gb

    def syndrome(self, c, hd_row):
        hd_row = np.ascontiguousarray(hd_row, dtype=np.uint8)
        return self.lib.oracle_syndrome(ctypes.byref(c), hd_row.ctypes.data)

    def setup_ldpc(self, A, rate):
        out = (ctypes.c_int * 10)()
        self.lib.oracle_setup_ldpc(A, rate, out)
        return list(out)

    def base_matrix(self, kind, ils):
        rows, cols = {1: (46, 68), 2: (42, 52), 3: (12, 24)}[kind]
        out = np.empty(rows * cols, dtype=np.int32)
        self.lib.oracle_base_matrix(kind, ils, out.ctypes.data)
        return out.reshape(rows, cols)


_ORACLE = None


def load():
    global _ORACLE
    if _ORACLE is None:
        if not os.path.exists(LIB):
            subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "-s"], check=True)
        _ORACLE = Oracle(ctypes.CDLL(LIB))
    return _ORACLE
