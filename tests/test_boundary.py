"""The C-ABI library loads and exports what include/ct_hip.h declares (no GPU).

Also checks the host-only entry point ct_phantom_params bit for bit against
the oracle's parameter draw (it runs without a device).
"""
import ctypes
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "ct_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(ct_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    import qtctreconstruction_amd as ct
    L = ct.load_library()
    names = _declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), f"{n} declared in include/ct_hip.h but not exported"
    # the Python host side binds exactly the declared set
    assert sorted(ct.exported_symbols()) == names


def test_phantom_params_match_oracle(O):
    import qtctreconstruction_amd as ct
    for seed, n in ((0, 1), (42, 24), (2**63 + 5, 200)):
        a = ct.phantom_params(seed, n)
        b = O.phantom_params(seed, n)
        for i in range(n):
            assert bytes(a[i]) == bytes(b[i]), (seed, i)


def test_no_device_is_reported_not_faked():
    """On a host without a GPU the backend says so instead of computing on the CPU."""
    import qtctreconstruction_amd as ct
    if ct.load_library().ct_hip_device_count() > 0:
        return
    h = ctypes.c_void_p()
    rc = ct.load_library().ct_hip_create(0, 0, ctypes.byref(h))
    assert rc == ct.CT_ERR_NODEVICE
    try:
        ct.HipBackend(0)
    except ct.BackendError as e:
        assert e.code == ct.CT_ERR_NODEVICE
    else:
        raise AssertionError("HipBackend constructed without a device")


def test_product_does_not_import_oracle():
    pkg = os.path.join(ROOT, "qtctreconstruction_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".cpp", ".hip", ".h")):
                txt = open(os.path.join(dirpath, f), errors="replace").read()
                assert "oracle" not in re.sub(r"#.*|//.*", "", txt).lower() or f == "__init__.py" \
                    and "import oracle" not in txt, f


def test_shape_limits_refused_without_a_device():
    """ct_hip_check_shape is the check every entry point makes before any
    device work: Radon edges beyond 2048 px or more than 4096 bins, and FFT
    filters longer than 4096 (bins > 1448), are refused; sizes <= 0 are invalid."""
    import qtctreconstruction_amd as ct
    R, F = ct.CT_STAGE_RADON, ct.CT_STAGE_FBP
    assert ct.check_shape(R, 2048, 2048, 4096) == ct.CT_OK
    assert ct.check_shape(R, 2049, 16, 16) == ct.CT_ERR_UNSUPPORTED
    assert ct.check_shape(R, 16, 2049, 16) == ct.CT_ERR_UNSUPPORTED
    assert ct.check_shape(R, 16, 16, 4097) == ct.CT_ERR_UNSUPPORTED
    assert ct.check_shape(R, 0, 16, 16) == ct.CT_ERR_INVALID
    assert ct.check_shape(F, 0, 0, 1448) == ct.CT_OK          # nfft 4096
    assert ct.check_shape(F, 0, 0, 1449) == ct.CT_ERR_UNSUPPORTED
    assert ct.check_shape(F, 0, 0, 1449, int(ct.FilterBackend.GEMM_SPLIT_BF16)) == ct.CT_OK
    assert ct.check_shape(F, 0, 0, 0) == ct.CT_ERR_INVALID
    assert ct.check_shape(7, 1, 1, 1) == ct.CT_ERR_INVALID
