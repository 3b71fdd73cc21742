"""The C-ABI library loads and exports every symbol include/fe_hip.h declares (CPU only, no compute)."""
import ctypes as ct
import re

from conftest import ROOT, load_package


def declared_symbols():
    text = (ROOT / "include" / "fe_hip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b([a-z_][a-z0-9_]*)\s*\(", text)) - {"if", "defined"})


def test_header_symbols_exported():
    fe = load_package()
    lib = ct.CDLL(str(fe.LIB_PATH))
    names = declared_symbols()
    assert len(names) >= 9
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding covers exactly the declared surface
    assert sorted(fe.exported_symbols()) == names


def test_package_refuses_without_library(tmp_path):
    import importlib.util
    import shutil
    import sys
    pkg = tmp_path / "pkg"
    shutil.copytree(ROOT / "3fold-field-engine_amd", pkg, ignore=shutil.ignore_patterns("*.so", "csrc"))
    spec = importlib.util.spec_from_file_location("fe_amd_nolib", pkg / "__init__.py",
                                                  submodule_search_locations=[str(pkg)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["fe_amd_nolib"] = mod
    try:
        import pytest
        with pytest.raises(ImportError):
            spec.loader.exec_module(mod)
    finally:
        sys.modules.pop("fe_amd_nolib", None)


def test_host_validation_rejects_without_gpu():
    """Invalid inputs are refused on the host with status 1 before any device call."""
    import numpy as np
    fe = load_package()
    lib = fe._lib.lib()
    for shape in ((2, 2, 2), (129, 2, 2)):
        m = np.zeros(shape, np.uint8)
        out = np.empty(shape, np.int32)
        assert lib.edt_squared(m.ctypes.data, *shape, out.ctypes.data) == 1
    m = np.full((2, 2, 2), 2, np.uint8)
    assert lib.edt_squared_large(m.ctypes.data, 2, 2, 2, out.ctypes.data) == 1
    assert lib.edt_pair_create(0, 4, 4) is None
    assert lib.edt_pair_create(513, 4, 4) is None
