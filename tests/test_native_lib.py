"""The C-ABI library loads and exports every symbol include/fvm_hip.h declares (no GPU)."""
import os
import re

import pytest

from fvm_solver_amd import _native as nat

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "fvm_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(fvm_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_the_bound_api():
    assert _declared() == sorted(nat.exported_symbols())


@pytest.mark.skipif(not os.path.exists(nat.LIB_PATH), reason="libfvm_hip.so not built")
def test_library_exports_every_declared_symbol():
    lib = nat.load_library()
    for name in _declared():
        assert hasattr(lib, name), name
    assert lib.fvm_abi_version() == nat.ABI_VERSION


@pytest.mark.skipif(not os.path.exists(nat.LIB_PATH), reason="libfvm_hip.so not built")
def test_argument_errors_are_reported_without_a_gpu():
    lib = nat.load_library()
    rc = lib.fvm_ell_spmm_f32(None, 0, 0, None, 0, 0, 0, -1, 0, None, 0, 0, 1, None, None, 0, 0, None)
    assert rc == 1 and b"invalid" in lib.fvm_last_error()
    with pytest.raises(RuntimeError):
        nat.check(rc, "fvm_ell_spmm_f32")


def test_channel_config_matches_golden_parameters():
    from fvm_solver_amd.config_fvm import ConfigChannel, U_INF
    from tests.golden_cases import load
    d = load("ref_2d.npz")
    cfg = ConfigChannel()
    for k, v in zip(d["phys_keys"], d["phys"]):
        assert abs(float(getattr(cfg, str(k))) - float(v)) <= 1e-12 * max(1.0, abs(float(v))), k
    assert abs(U_INF - float(d["u_inf"])) < 1e-12
