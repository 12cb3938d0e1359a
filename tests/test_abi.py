"""The C ABI library loads, exports every symbol include/aerhip.h declares,
and refuses to run without a HIP device (no silent CPU fallback)."""
import ctypes
import pathlib
import re

import pytest

from qiskit_aer_amd import _lib as L

HEADER = pathlib.Path(__file__).resolve().parent.parent / "include" / "aerhip.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(aerhip_[a-z0-9_]+)\s*\(", text)))


def test_library_is_built():
    assert L.LIB_PATH.exists(), "run __graft_entry__.build() first"


def test_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(L.LIB_PATH))
    names = declared_functions()
    assert len(names) >= 35
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    assert set(declared_functions()) <= set(L.exported_symbols())


def test_struct_layouts_match_header():
    # aerhip_op: int32, uint32, ptr, ptr, uint32, uint32, ptr, ptr, ptr
    assert ctypes.sizeof(L.AerhipOp) == 4 + 4 + 8 + 8 + 4 + 4 + 8 + 8 + 8
    assert ctypes.sizeof(L.AerhipRunConfig) == 4 + 4 + 8 + 8 + 4 * 4 + 8


def test_defaults_match_reference():
    cfg = L.AerhipRunConfig()
    L.lib().aerhip_run_config_default(ctypes.byref(cfg))
    assert (cfg.fusion_max_qubit, cfg.fusion_threshold, cfg.sample_index_size) == (5, 14, 10)
    assert L.lib().aerhip_required_memory_mb(30, 0) == 16384
    assert L.lib().aerhip_required_memory_mb(30, 1) == 8192


def test_no_device_means_error_not_fallback():
    if L.lib().aerhip_device_count() > 0:
        pytest.skip("a HIP device is present")
    from qiskit_aer_amd import QubitVector
    with pytest.raises(L.AerhipError):
        QubitVector(4)


def _build_cpp(tmp_path, src):
    import shutil
    import subprocess
    if not shutil.which("g++"):
        pytest.skip("g++ not available")
    root = HEADER.parent.parent
    out = tmp_path / "a.out"
    return subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", f"-I{root / 'include'}", str(src),
                           f"-L{L.LIB_PATH.parent}", "-laerhip", f"-Wl,-rpath,{L.LIB_PATH.parent}",
                           "-o", str(out)], capture_output=True, text=True)


def test_cpp_adaptor_compiles_and_links(tmp_path):
    """include/qubitvector_hip.hpp (the QubitVector drop-in) builds against the ABI."""
    src = tmp_path / "a.cpp"
    src.write_text('#include "qubitvector_hip.hpp"\n'
                   "template class aerhip::QubitVectorHip<double>;\n"
                   "template class aerhip::QubitVectorHip<float>;\n"
                   "int main() { return 0; }\n")
    r = _build_cpp(tmp_path, src)
    assert r.returncode == 0, r.stderr


def test_cpp_adaptor_takes_every_state_call(tmp_path):
    """Every call Statevector::State makes on its qreg_ (statevector_state.hpp),
    with the reference's argument types (QV::Rotation enum, AER::Vector
    results, reg_t / cvector_t arguments), compiles against the adaptor and
    links against libaerhip.so (tests/adaptor/state_calls.cpp)."""
    src = pathlib.Path(__file__).resolve().parent / "adaptor" / "state_calls.cpp"
    r = _build_cpp(tmp_path, src)
    assert r.returncode == 0, r.stderr
