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


HEADER_STRUCTS = ("bepu_settings", "bepu_motion_state", "bepu_material", "bepu_contact", "bepu_manifold",
                  "bepu_impulses", "bepu_step_stats", "bepu_body_desc", "bepu_tuning")


@pytest.fixture(scope="module")
def header_sizes(tmp_path_factory):
    """sizeof of every header struct as gcc lays it out."""
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    d = tmp_path_factory.mktemp("sz")
    src = d / "sz.c"
    body = " ".join('printf("%%zu ", sizeof(%s));' % n for n in HEADER_STRUCTS)
    src.write_text('#include <stdio.h>\n#include "%s/include/bepu_gpu.h"\nint main(){%s return 0;}' % (ROOT, body))
    exe = d / "sz"
    subprocess.run([cc, str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    return dict(zip(HEADER_STRUCTS, (int(x) for x in out)))


def test_struct_sizes_match_header_layout(header_sizes):
    """ctypes layouts equal the C compiler's layout of the header structs."""
    from bepuphysics1_amd import _lib
    mine = dict(zip(HEADER_STRUCTS, (ctypes.sizeof(t) for t in (
        _lib.Settings, _lib.MotionState, _lib.Material, _lib.Contact, _lib.Manifold, _lib.Impulses,
        _lib.StepStats, _lib.BodyDesc, _lib.Tuning))))
    assert mine == header_sizes


def test_oracle_binding_layout_matches_its_library(header_sizes):
    """The oracle binding's structs have the sizes the oracle library was
    compiled with (it fills the header's structs: a shorter ctypes struct
    would be written past its end)."""
    from oracle import refcpu
    L = refcpu.lib()
    types = (refcpu.Settings, refcpu.MotionState, refcpu.Material, refcpu.Contact, refcpu.Manifold,
             refcpu.Impulses, refcpu.StepStats)
    for i, t in enumerate(types):
        assert ctypes.sizeof(t) == L.refcpu_sizeof(i) == header_sizes[HEADER_STRUCTS[i]], t.__name__


_CS_SIZES = {"float": 4, "int": 4, "uint": 4, "ulong": 8, "long": 8, "IntPtr": 8}


def _csharp_struct_sizes(text):
    """Sequential-layout sizes of the structs of a C# source (scalar fields
    and nested structs declared earlier in it)."""
    sizes, aligns = {}, {}
    for name, body in re.findall(r"public struct (\w+)\s*(?://[^\n]*)?\s*\{(.*?)\}", text, flags=re.S):
        off, al = 0, 1
        for typ, names in re.findall(r"public\s+(\w+)\s+([\w\s,]+);", body):
            sz = _CS_SIZES.get(typ, sizes.get(typ))
            a = min(_CS_SIZES.get(typ, aligns.get(typ, 4)), 8)
            assert sz is not None, (name, typ)
            for _ in names.split(","):
                off = (off + a - 1) // a * a + sz
                al = max(al, a)
        sizes[name] = (off + al - 1) // al * al
        aligns[name] = al
    return sizes


def test_csharp_stub_layout_matches_header(header_sizes):
    """The P/Invoke stub's structs in INTEGRATION.md have the header's sizes,
    so bepu_get_stats and the others write exactly into them."""
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    cs = _csharp_struct_sizes(text)
    pairs = {"BepuSettings": "bepu_settings", "BepuMotionState": "bepu_motion_state",
             "BepuMaterial": "bepu_material", "BepuStepStats": "bepu_step_stats", "BepuTuning": "bepu_tuning",
             "BepuBodyDesc": "bepu_body_desc"}
    for c, h in pairs.items():
        assert cs.get(c) == header_sizes[h], (c, cs.get(c), header_sizes[h])


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
