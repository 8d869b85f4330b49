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


@pytest.mark.skipif(not os.path.exists(nat.LIB_PATH), reason="libfvm_hip.so not built")
def test_tile_list_arguments_are_checked_without_a_gpu():
    """A stage's tile list needs the tile-fused stage and 1 <= n_tile_list <= ceil(N / 256);
    both are rejected before anything touches the device (fake device pointers)."""
    import ctypes
    lib = nat.load_library()
    fake = ctypes.c_void_p(0x10000)
    m = nat.FvmMeshDev(dim=2, n_cells=1000, n_facets=2000, n_meshes=1, fac_cell=fake, fac_geo=fake)
    p = nat.FvmPhys(lim_p=4)
    w = nat.FvmWork(gp=fake, fflux=fake)
    st = nat.FvmStage(prim=fake, U0=fake, acc=fake, prim_next=fake, U_out=fake, dt=fake, tile_list=fake,
                      n_tile_list=2)
    rc = lib.fvm_rk_stage_f32(ctypes.byref(m), ctypes.byref(p), ctypes.byref(w), ctypes.byref(st), None)
    assert rc == 1 and b"tile list needs the tile-fused stage" in lib.fvm_last_error()
    m.fz_tile = fake
    for n in (0, 5):                                    # ceil(1000 / 256) = 4
        st.n_tile_list = n
        rc = lib.fvm_rk_stage_f32(ctypes.byref(m), ctypes.byref(p), ctypes.byref(w), ctypes.byref(st), None)
        assert rc == 1 and f"n_tile_list {n} outside [1, 4]".encode() in lib.fvm_last_error(), n


def test_channel_config_matches_golden_parameters():
    from fvm_solver_amd.config_fvm import ConfigChannel, U_INF
    from tests.golden_cases import load
    d = load("ref_2d.npz")
    cfg = ConfigChannel()
    for k, v in zip(d["phys_keys"], d["phys"]):
        assert abs(float(getattr(cfg, str(k))) - float(v)) <= 1e-12 * max(1.0, abs(float(v))), k
    assert abs(U_INF - float(d["u_inf"])) < 1e-12
