"""C ABI checks that need no GPU: the library loads and exports every entry
point include/bepu_gpu.h declares; the Python binding covers them all; host
logic (argument validation before any device work) behaves as documented."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for fn in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names += re.findall(r"\b(bepu_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "bepu_world_create" in names and "bepu_step" in names
    assert len(names) >= 14


def test_library_exports_every_declared_symbol():
    from bepuphysics1_amd import _lib
    L = _lib.load()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_binding_covers_header():
    from bepuphysics1_amd import _lib
    assert set(declared_functions()) <= set(_lib.EXPORTS)


def test_struct_sizes_match_header_layout(tmp_path):
    """ctypes layouts equal the C compiler's layout of the header structs."""
    import shutil
    import subprocess
    from bepuphysics1_amd import _lib
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    src = tmp_path / "sz.c"
    src.write_text('#include <stdio.h>\n#include "%s/include/bepu_gpu.h"\n'
                   'int main(){printf("%%zu %%zu %%zu %%zu %%zu %%zu %%zu %%zu %%zu", sizeof(bepu_settings),'
                   'sizeof(bepu_motion_state), sizeof(bepu_material), sizeof(bepu_contact), sizeof(bepu_manifold),'
                   'sizeof(bepu_impulses), sizeof(bepu_step_stats), sizeof(bepu_body_desc), sizeof(bepu_tuning));}'
                   % ROOT)
    exe = tmp_path / "sz"
    subprocess.run([cc, str(src), "-o", str(exe)], check=True)
    sizes = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    mine = [ctypes.sizeof(t) for t in (_lib.Settings, _lib.MotionState, _lib.Material, _lib.Contact,
                                       _lib.Manifold, _lib.Impulses, _lib.StepStats, _lib.BodyDesc, _lib.Tuning)]
    assert mine == sizes


def test_defaults_match_reference_constants():
    from bepuphysics1_amd import _lib
    s = _lib.default_settings()
    assert s.iteration_limit == 10 and s.min_iteration_count == 1
    assert s.min_impulse == pytest.approx(0.001)
    assert s.allowed_penetration == pytest.approx(0.01)
    assert s.max_contact_distance == pytest.approx(0.1)
    assert s.penetration_recovery_stiffness == pytest.approx(0.2)
    assert s.max_penetration_recovery_speed == pytest.approx(2.0)
    assert s.softness == pytest.approx(0.001)
    assert s.static_friction_velocity_threshold == pytest.approx(0.2)
    assert s.linear_damping == pytest.approx(0.03) and s.angular_damping == pytest.approx(0.15)
    assert s.dt == pytest.approx(1 / 60)


def test_create_without_device_fails_loudly():
    """No silent CPU fallback: without a HIP device the world cannot exist."""
    from bepuphysics1_amd import _lib
    L = _lib.load()
    if L.bepu_device_available():
        pytest.skip("a HIP device is present")
    h = ctypes.c_void_p()
    st = L.bepu_world_create(None, 16, ctypes.byref(h))
    assert st == 5  # BEPU_E_DEVICE
    from bepuphysics1_amd import Space, Box
    sp = Space(capacity=4)
    with pytest.raises(_lib.BepuError):
        sp.Add(Box((0, 0, 0), 1, 1, 1, 1))


def test_invalid_mass_rejected_on_host():
    from bepuphysics1_amd import Box
    with pytest.raises(ValueError):
        Box((0, 0, 0), 1, 1, 1, float("nan"))


def test_bad_args_rejected():
    from bepuphysics1_amd import _lib
    L = _lib.load()
    assert L.bepu_world_create(None, 0, ctypes.byref(ctypes.c_void_p())) == 1
    assert L.bepu_step(None, 1) == 1
    assert L.bepu_world_destroy(None) == 1
