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


# ---- the bindings against the header: field names, order and types, field ids, ABI version ----
JULIA = os.path.join(ROOT, "fastisostasy.jl_amd", "julia", "FastIsostasyHIP.jl")
C_TO_JULIA = {"int32_t": "Int32", "int64_t": "Int64", "double": "Float64"}
C_TO_CTYPES = {"int32_t": ctypes.c_int32, "int64_t": ctypes.c_int64, "uint64_t": ctypes.c_uint64,
               "double": ctypes.c_double}


def header_text():
    return re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)


def header_struct(name):
    """[(field, C type, array length or 0)] of `typedef struct name { ... } name;`."""
    src = header_text()
    body = re.search(r"typedef struct %s \{(.*?)\} %s;" % (name, name), src, flags=re.S).group(1)
    macros = dict((k, int(v)) for k, v in re.findall(r"#define (\w+) (\d+)", src))
    out = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl:
            continue
        ctype, rest = decl.split(None, 1)
        for item in rest.split(","):
            item = item.strip()
            m = re.match(r"(\w+)\s*(?:\[(\w+)\])?$", item)
            n = m.group(2)
            out.append((m.group(1), ctype, (int(n) if n.isdigit() else macros[n]) if n else 0))
    return out


def julia_struct(name):
    src = open(JULIA).read()
    body = re.search(r"(?:mutable )?struct %s\n(.*?)\nend" % name, src, flags=re.S).group(1)
    return [(f, t) for f, t in re.findall(r"(\w+)::(\w+)", body)]


def norm(name):
    return name.replace("_", "").lower()


def header_field_ids():
    src = header_text()
    block = re.search(r"enum \{([^}]*FI_LITHO_RIGIDITY[^}]*)\}", src, flags=re.S).group(1)
    return {norm(k): int(v) for k, v in re.findall(r"FI_(\w+)\s*=\s*(\d+)", block)}


@pytest.mark.parametrize("cname,jname", [("fi_config", "FiConfig"), ("fi_stats", "FiStats")])
def test_julia_structs_match_header(cname, jname):
    h = header_struct(cname)
    assert all(n == 0 for _, _, n in h)
    assert julia_struct(jname) == [(f, C_TO_JULIA[t]) for f, t, _ in h]


@pytest.mark.parametrize("cname,cls", [("fi_config", "FiConfig"), ("fi_stats", "FiStats"),
                                       ("fi_batch_info", "FiBatchInfo"), ("fi_rkc_info", "FiRkcInfo"),
                                       ("fi_tune_info", "FiTuneInfo")])
def test_ctypes_structs_match_header(cname, cls):
    fields = getattr(_abi, cls)._fields_
    h = header_struct(cname)
    assert [f for f, _ in fields] == [f for f, _, _ in h]
    for (f, ct), (_, t, n) in zip(fields, h):
        assert ct == (C_TO_CTYPES[t] * n if n else C_TO_CTYPES[t]), f


def test_field_ids_match_header(ref):
    ids = header_field_ids()
    julia = dict(re.findall(r"(\w+) = (\d+)", re.search(r"const FI = \((.*?)\)\n", open(JULIA).read(),
                                                        flags=re.S).group(1)))
    for table in ({norm(k): int(v) for k, v in julia.items()}, {norm(k): v for k, v in _abi.FIELDS.items()},
                  {norm(k): v for k, v in ref.FIELDS.items()}):
        assert table == ids


def test_abi_version_matches_bindings():
    v = int(re.search(r"#define FI_ABI_VERSION (\d+)", open(HEADER).read()).group(1))
    assert _abi.FI_ABI_VERSION == v
    assert int(re.search(r"const FI_ABI_VERSION = (\d+)", open(JULIA).read()).group(1)) == v


def test_engine_loader_refuses_another_engine(ref):
    """The oracle exports the same ABI; loading it as the engine is refused unless asked for."""
    with pytest.raises(RuntimeError, match="cpu-oracle"):
        _abi.load_engine(ref.LIB_PATH)
    lib = _abi.load_engine(ref.LIB_PATH, allow_other_engine=True)
    assert _abi.engine_identity(lib) == ("cpu-oracle-fp64", os.path.abspath(ref.LIB_PATH))


def test_oracle_ctypes_stats_match_header(ref):
    h = header_struct("fi_stats")
    assert [f for f, _ in ref.RefStats._fields_] == [f for f, _, _ in h]
    assert [ct for _, ct in ref.RefStats._fields_] == [C_TO_CTYPES[t] for _, t, _ in h]
