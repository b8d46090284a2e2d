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
