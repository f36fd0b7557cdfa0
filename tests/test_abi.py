"""The C-ABI library loads, exports every entry point include/mcpt.h
declares, and fails loudly (no CPU fallback) without a gfx950 device."""
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    names = set()
    for h in ("mcpt.h",):
        with open(os.path.join(ROOT, "include", h)) as f:
            text = f.read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(mcpt_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_every_declared_symbol_is_exported(mcpt):
    names = header_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(mcpt._lib, n), n
    assert set(names) <= set(mcpt.EXPORTS), set(names) - set(mcpt.EXPORTS)


def test_record_layouts(mcpt):
    import ctypes as C
    # sizes printed by gcc for the structs of include/mcpt.h
    assert C.sizeof(mcpt.CameraDesc) == 72 and C.sizeof(mcpt.Camera) == 112
    assert C.sizeof(mcpt.LbvhDump) == 104 and C.sizeof(mcpt.RenderOpts) == 40
    assert C.sizeof(mcpt.RenderStats) == 96 and C.sizeof(mcpt.EnvDesc) == 24
    assert C.sizeof(mcpt.ServerOpts) == 32 and C.sizeof(mcpt.Job) == 136
    assert C.sizeof(mcpt.JobResult) == 48 and C.sizeof(mcpt.ServerStats) == 88
    assert C.sizeof(mcpt.CacheOpts) == 28 and C.sizeof(mcpt.CacheStats) == 72
    assert C.sizeof(mcpt.BatchTune) == 16
    assert C.sizeof(mcpt.ObjOpts) == 72 and C.sizeof(mcpt.ObjStats) == 32
    assert C.sizeof(mcpt.TextureDesc) == 48 and mcpt.MAT_DTYPE.itemsize == 48
    assert C.sizeof(mcpt.DensityGrid) == 48
    assert mcpt.HOT_DTYPE.itemsize == 36 and mcpt.COLD_DTYPE.itemsize == 64


def test_generators_deterministic(mcpt):
    a = mcpt.gen_scene(mcpt.GEN_SOUP, seed=7, param=1000)
    b = mcpt.gen_scene(mcpt.GEN_SOUP, seed=7, param=1000)
    c = mcpt.gen_scene(mcpt.GEN_SOUP, seed=8, param=1000)
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
    assert not np.array_equal(a[0], c[0])
    assert len(a[0]) == 1016  # 1000 + 16 emitters
    hot, cold, mats = mcpt.gen_scene(mcpt.GEN_CORNELL)
    assert len(hot) == 36 and (mats["type"][cold["material"]] == mcpt.MAT_EMISSIVE).sum() == 2
    # Cornell albedos are sRGB-decoded at construction (make_solid_texture)
    assert np.allclose(mats["albedo"][0], (0.73 ** 2.4 if False else ((0.73 + 0.055) / 1.055) ** 2.4), rtol=1e-5)
    sky = mcpt.gen_sky(512, 256)
    assert sky.shape == (256, 512, 4) and np.all(sky[..., 3] == 1)
    # ground colour below the horizon (make_sky.py: 0.8 * (0.25, 0.22, 0.18) plus the glow)
    assert abs(sky[250, 0, 0] - 0.2) < 0.01


def test_bad_arguments_are_errors(mcpt):
    with pytest.raises(mcpt.McptError):
        mcpt.gen_scene(99)
    d = mcpt.gen_camera(mcpt.GEN_CORNELL, 100, 100)
    d.height = 0
    with pytest.raises(mcpt.McptError):
        mcpt.camera_init(d)


def test_no_device_fails_loudly(mcpt):
    import ctypes as C
    n = C.c_int(0)
    hip = C.CDLL("libamdhip64.so")
    if hip.hipGetDeviceCount(C.byref(n)) == 0 and n.value > 0:
        pytest.skip("a GPU is present")
    with pytest.raises(mcpt.McptError) as e:
        mcpt.Context()
    assert e.value.code == mcpt.ERR_NODEVICE
