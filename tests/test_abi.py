"""The C ABI: both libraries load and export every entry point include/fastiso.h declares, and
the ctypes structures match the C layout. No compute calls (runs without a GPU)."""
import ctypes
import os
import re
import subprocess

import pytest

from fastiso import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fastiso.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(fi_[a-z_0-9]+)\s*\(", src)) - {"fi_output_cb"})


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("fi_create", "fi_rhs", "fi_integrate", "fi_sparse_update", "fi_integrator_init", "fi_get_field"):
        assert must in names


def test_hip_engine_exports_all_symbols():
    if not os.path.exists(_abi.ENGINE_PATH):
        pytest.skip("libfastiso_hip.so not built")
    lib = ctypes.CDLL(_abi.ENGINE_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    _abi.declare(lib)
    assert lib.fi_abi_version() == _abi.FI_ABI_VERSION
    assert lib.fi_engine_name() == b"hip-gfx950"


def test_oracle_exports_all_symbols(ref):
    lib = ref.lib()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_struct_layout_matches_header(tmp_path):
    prog = tmp_path / "layout.c"
    prog.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "fastiso.h"\n'
                    'int main(void){printf("%zu %zu %zu %zu %zu %zu %zu\\n", sizeof(fi_config), sizeof(fi_stats),'
                    ' offsetof(fi_config, dt_sparse), offsetof(fi_config, device), offsetof(fi_stats, last_EEst),'
                    ' sizeof(fi_batch_info), offsetof(fi_batch_info, cell_attempts));'
                    'return 0;}\n')
    exe = tmp_path / "layout"
    subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), str(prog), "-o", str(exe)])
    out = [int(x) for x in subprocess.check_output([str(exe)]).split()]
    assert out == [ctypes.sizeof(_abi.FiConfig), ctypes.sizeof(_abi.FiStats), _abi.FiConfig.dt_sparse.offset,
                   _abi.FiConfig.device.offset, _abi.FiStats.last_EEst.offset, ctypes.sizeof(_abi.FiBatchInfo),
                   _abi.FiBatchInfo.cell_attempts.offset]


def test_default_config_matches_reference_defaults(ref):
    """fi_default_config of both libraries = PhysicalConstants() and SolidEarth() defaults."""
    cfg = _abi.FiConfig()
    L = ref.lib()
    L.fi_default_config.argtypes = [ctypes.POINTER(_abi.FiConfig), ctypes.c_int32, ctypes.c_int32,
                                    ctypes.c_double, ctypes.c_double]
    L.fi_default_config(ctypes.byref(cfg), 64, 64, 3e6, 3e6)
    assert cfg.g == 9.8 and cfg.rho_ice == 910 and cfg.rho_seawater == 1023 and cfg.rho_uppermantle == 3400
    assert cfg.rho_litho == 3200 and cfg.litho_poissonratio == 0.28 and cfg.seconds_per_year == 31557600.0
    assert cfg.reltol == 1e-5 and cfg.abstol == 1e-6 and cfg.dt_sparse == 10.0


def test_engine_loader_fails_loudly(tmp_path):
    with pytest.raises(RuntimeError):
        _abi.load_engine(str(tmp_path / "missing.so"))
