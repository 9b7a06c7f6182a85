"""The C-ABI boundary without a GPU: libvedo_hip.so loads and exports every
function include/vedo_gpu.h declares, the ctypes binding knows all of them,
and the test oracle library exports what oracle/vedo_oracle.h declares."""
import ctypes as C
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header, prefix):
    text = open(os.path.join(ROOT, header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(" + prefix + r"\w+)\s*\(", text)))


def test_hip_library_exports_every_declared_entry_point():
    names = _declared("include/vedo_gpu.h", "vg_")
    assert len(names) >= 20
    lib = C.CDLL(os.path.join(ROOT, "vedo_amd", "libvedo_hip.so"))
    missing = [s for s in names if not hasattr(lib, s)]
    assert not missing, missing
    from vedo_amd.gpu import EXPORTS
    assert sorted(EXPORTS) == names


def test_create_without_device_fails_loudly():
    """Without a GPU, vg_create fails with an error message (no CPU fallback)."""
    import pytest
    torch = pytest.importorskip("torch")
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from vedo_amd import scenes
    from vedo_amd.gpu import GpuError, GpuSimulation
    wa = scenes.box_scene(8, seed=1, spacing=3.0).to_arrays(up_iact=5, record=100)
    with pytest.raises(GpuError, match="device"):
        GpuSimulation(wa)


def test_oracle_library_exports_every_declared_function():
    names = _declared("oracle/vedo_oracle.h", "vo_")
    lib = C.CDLL(os.path.join(ROOT, "oracle", "libvedo_oracle.so"))
    missing = [s for s in names if not hasattr(lib, s)]
    assert not missing, missing
