"""libfmd.so: it loads and exports every function include/fmd.h declares.

No compute call is made here (no GPU in the CPU suite)."""
import ctypes as C
import subprocess

from fuller_md_en_amd import _native


def test_header_declares_entry_points():
    syms = _native.declared_symbols()
    for s in ("fmd_rigid_forces", "fmd_rigid_step_npt", "fmd_rigid_nlist", "fmd_engine_run",
              "fmd_last_error"):
        assert s in syms


def test_library_exports_every_declared_symbol(fmd_lib):
    out = subprocess.run(["nm", "-D", "--defined-only", _native.LIB_PATH], check=True,
                         capture_output=True, text=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    missing = [s for s in _native.declared_symbols() if s not in exported]
    assert not missing, missing


def test_pure_host_helpers(fmd_lib):
    p = _native.LJParams()
    fmd_lib.fmd_lj_params_rigid(C.byref(p))
    assert abs(p.rcut - 3 * 3.431) < 1e-15 and p.rmin2 == 0.25
    assert abs(p.vshift - 4 * 2.635e-3 * ((1 / 3) ** 12 - (1 / 3) ** 6)) < 1e-18
    s = _native.NPTState()
    fmd_lib.fmd_make_npt(C.byref(s), 300.0, 0.0, 6 * 108 - 3)
    kB = 8.617333262e-5
    assert abs(s.Q - 645 * kB * 300 * 1e4) < 1e-9 * s.Q
    assert abs(s.W - 654 * kB * 300 * 1e6) < 1e-9 * s.W
    assert fmd_lib.fmd_version() >= 100
