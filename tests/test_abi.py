"""The C-ABI library loads, exports every entry point include/*.h declares, and
its host-side checks (parameter validation, shape limits) answer with the
reference's messages.  CPU only: no compute call reaches a device."""
from __future__ import annotations

import ctypes as C
import glob
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "inspire-rs_amd", "libinspire_mi355x.so")


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "inspire-rs_amd"), "-j8"], check=True)
    import inspire_mi355x as M

    return M.load_library(LIB)


def declared_functions():
    names = []
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"^\s*[A-Za-z_][\w\s\*]*?\b(inspire_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_the_respond_api():
    names = declared_functions()
    for must in ["inspire_respond", "inspire_respond_seeded", "inspire_respond_inspiring",
                 "inspire_respond_seeded_inspiring", "inspire_respond_with_variant", "inspire_respond_batch",
                 "inspire_ctx_create", "inspire_db_load", "inspire_ntt_forward", "inspire_external_product"]:
        assert must in names


def test_every_declared_symbol_is_exported(lib):
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_every_symbol():
    import inspire_mi355x as M

    bound = {name for name, _, _ in M.EXPORTS}
    assert set(declared_functions()) == bound


def _params(M, **kw):
    p = M.InspireParams(**kw)
    return p._c()


def _create(lib, M, p):
    h = C.c_void_p()
    seed = (C.c_uint8 * 32)()
    rc = lib.inspire_ctx_create(C.byref(p), seed, C.byref(h))
    msg = lib.inspire_last_error().decode()
    if h:
        lib.inspire_ctx_destroy(h)
    return rc, msg


@pytest.mark.parametrize("kw,msg", [
    (dict(ring_dim=1000), "ring_dim must be a power of two"),
    (dict(crt_moduli=(268369921, 249561091)), "CRT moduli must be ≡ 1 (mod 2d) for NTT"),
    (dict(crt_moduli=(268369921, 268369921)), "CRT moduli must be coprime"),
])
def test_invalid_params_use_reference_messages(lib, kw, msg):  # params.rs:348-387
    import inspire_mi355x as M

    rc, m = _create(lib, M, _params(M, **kw))
    assert rc == M.ERR_INVALID_PARAMS and m == msg


def test_q_must_equal_product(lib):
    import inspire_mi355x as M

    p = _params(M)
    p.q = 12345
    rc, m = _create(lib, M, p)
    assert rc == M.ERR_INVALID_PARAMS and m == "q must equal product of CRT moduli"


def test_unsupported_shapes_are_refused(lib):
    import inspire_mi355x as M

    # the reference tests' 60-bit single modulus is valid but not on the 32-bit path
    rc, m = _create(lib, M, _params(M, ring_dim=256, crt_moduli=(1152921504606830593,)))
    assert rc == M.ERR_UNSUPPORTED and "2^29" in m
    rc, m = _create(lib, M, _params(M, gadget_base=3 << 18))
    assert rc == M.ERR_UNSUPPORTED


def test_null_arguments(lib):
    import inspire_mi355x as M

    assert lib.inspire_ctx_create(None, None, None) == M.ERR_INVALID_ARGUMENT
    assert lib.inspire_respond(None, None, None) == M.ERR_INVALID_ARGUMENT
    lib.inspire_ctx_destroy(None)  # no-op


def test_binding_fails_loudly_without_the_library(tmp_path):
    import inspire_mi355x as M

    with pytest.raises(RuntimeError, match="no CPU fallback"):
        M.load_library(str(tmp_path / "missing.so"))
