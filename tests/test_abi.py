"""CPU-side checks of the C ABI library and the host code (no GPU needed)."""
import ctypes
import os
import re

import numpy as np

import piper2_amd as pg
from piper2_amd import scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "piper_hip.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|uint32_t|const char\*)\s+(pg_\w+)\s*\(", src, re.M)))


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(pg.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(pg.SIGNATURES), set(names) ^ set(pg.SIGNATURES)


def test_struct_layouts_match_header(tmp_path):
    """The ctypes mirrors have the header's sizes and field offsets (a field the Python side lacks
    would make the library read past the caller's struct)."""
    import subprocess
    pairs = [("pg_scene_desc", pg.SceneDesc), ("pg_texture", pg.Texture), ("pg_material", pg.Material),
             ("pg_quad", pg.Quad), ("pg_camera", pg.Camera), ("pg_render_params", pg.RenderParams),
             ("pg_stats", pg.Stats)]
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "piper_hip.h"', "int main(void) {"]
    for cname, cls in pairs:
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for f in cls._fields_:
            lines.append(f'printf("{cname}.{f[0]} %zu\\n", offsetof({cname}, {f[0]}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = dict(line.split() for line in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.splitlines())
    for cname, cls in pairs:
        assert int(got[cname]) == ctypes.sizeof(cls), cname
        for f in cls._fields_:
            assert int(got[f"{cname}.{f[0]}"]) == getattr(cls, f[0]).offset, (cname, f[0])


def test_sobol_matrices_from_library_match_golden():
    g = np.load(os.path.join(ROOT, "tests", "golden", "sobol_columns.npz"))
    assert np.array_equal(pg.sobol_matrices(), g["columns"])


def test_device_create_reports_errors_without_throwing():
    h = ctypes.c_void_p()
    rc = pg.lib().pg_device_create(4096, ctypes.byref(h))  # no such device anywhere
    assert rc in (pg.PG_ERR_INVALID, pg.PG_ERR_DEVICE)
    assert pg.lib().pg_last_error()


def test_scene_generators_are_deterministic_and_sized():
    a, b = scenes.cornell(), scenes.cornell()
    for k in ("vertices", "indices", "materials", "quads", "camera"):
        assert np.array_equal(a[k], b[k])
    assert a["indices"].shape[0] == 10 + 24
    t = scenes.terrain_spheres()
    assert t["indices"].shape[0] == 2 * 999 * 499 + 200 * 1280
    assert t["quads"].shape[0] == 16
    mats = t["materials"]
    cond = mats[:, 0] == scenes.CONDUCTOR
    assert 60 < cond.sum() < 140
    assert np.all(mats[cond, 4] >= 0.05**2 - 1e-7) and np.all(mats[cond, 4] <= 0.6**2 + 1e-7)


def test_render_rect_fill_matches_reference_formula():
    ndc, rect = scenes.calc_render_rect(1920, 1080, 1920 / 1080)
    assert rect == (0, 0, 1920, 1080)
    assert np.allclose(ndc, (0.0, 0.0, 1.0, 1.0), atol=1e-6)
    ndc, _ = scenes.calc_render_rect(200, 100, 1.0)  # image wider than the sensor: crop rows
    assert np.isclose(ndc[2], 1.0) and np.isclose(ndc[3], 0.5) and np.isclose(ndc[1], 0.25)
