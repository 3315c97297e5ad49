"""The C++ component pair on the host: it builds, links the C ABI, and reports errors (no device
here: HIPAccelerator's pg_device_create fails with a message instead of aborting)."""
import ctypes
import os
import struct

import numpy as np
import pytest

from piper2_amd import scenes
from plugin_io import DRIVER, run_driver, write_scene


def test_driver_reports_errors_without_aborting(tmp_path):
    assert os.path.exists(DRIVER), "run __graft_entry__.build() first"
    r = run_driver()
    assert r.returncode == 2 and "usage" in r.stderr
    r = run_driver(str(tmp_path / "missing.pgsc"), str(tmp_path / "out"))
    assert r.returncode == 3 and "cannot open" in r.stderr
    bad = tmp_path / "bad.pgsc"
    bad.write_bytes(b"XXXX\x01\x00\x00\x00")
    r = run_driver(str(bad), str(tmp_path / "out"))
    assert r.returncode == 3 and "not a scene file" in r.stderr
    sc = scenes.cornell(width=16, height=16, spp=1)
    path = tmp_path / "cornell.pgsc"
    write_scene(sc, path)
    r = run_driver(str(path), str(tmp_path / "out"))
    if r.returncode == 0:
        pytest.skip("a GPU is present: the device path is covered by test_gpu_plugin.py")
    assert r.returncode == 3 and "pg_device_create" in r.stderr, (r.returncode, r.stderr)


PLUGIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "piper2_amd",
                      "libpiper_hip_plugin.so")


def test_module_entry_points():
    """piperInitModule / piperGetProtocol (Module.hpp:54-70) are exported; a host that dlopens the
    library gets protocol "GCC@STD@0" and the classes "Accelerator" and "Tracer"."""
    lib = ctypes.CDLL(PLUGIN)
    lib.piperGetProtocol.restype = ctypes.c_char_p
    assert lib.piperGetProtocol() == b"GCC@STD@0"
    assert hasattr(lib, "piperInitModule")
    r = run_driver("--module")
    assert r.returncode == 0, r.stdout + r.stderr
    assert "protocol GCC@STD@0" in r.stdout
    assert "accelerator host (no device) formats Native" in r.stdout
    assert "tracer HIPTracer" in r.stdout
    assert "unknown class: raised: module" in r.stdout and "has no class Denoiser" in r.stdout


def test_native_kernel_compile_and_launch():
    """compileKernel over a "Native" table (entry through a static redirect, one dynamic symbol),
    launchKernel with trivially copyable arguments and a resource look-up table, one context."""
    r = run_driver("--kernel")
    assert r.returncode == 0 and "kernel ok contexts 1" in r.stdout, r.stdout + r.stderr


def _payloads():
    r = run_driver("--payloads")
    assert r.returncode == 0, r.stdout + r.stderr
    out, calls = {}, {}
    for line in r.stdout.splitlines():
        f = line.split()
        if f[0] == "call":
            calls[int(f[1])] = (f[2], bytes.fromhex(f[3]))
        elif f[0] == "sampler":
            out["sampler"] = (int(f[1]), int(f[2]), bytes.fromhex(f[3]))
        elif f[0] != "resources":
            out[f[0]] = (f[1], bytes.fromhex(f[2]))
    return out, calls


TRAVERSAL = 0x1122334455667788


def test_payload_layouts_match_reference_structs():
    """Every built-in component packs the reference's own payload struct, read back here at the
    reference's 64-bit offsets (each component's Shared.hpp): texture inputs are call handles of
    ConstantData / CheckBoardData calls, the area light's sample is a planeSample call with CDFData,
    the sensor's PCData carries its traversal handle, and so on."""
    p, calls = _payloads()

    def call(handle, symbol):
        assert handle in calls, handle
        assert calls[handle][0] == symbol, (handle, calls[handle][0])
        return calls[handle][1]

    def constant(payload):  # ConstantData{uint32 channel, float value[4]}
        assert len(payload) == 20
        ch = struct.unpack_from("<I", payload)[0]
        return ch, np.frombuffer(payload, np.float32, 4, 4)

    sym, pay = p["matte"]  # MatteData{Call diffuse, Call roughness}
    assert sym == "matteInit" and len(pay) == 16
    d, r = struct.unpack("<QQ", pay)
    ch, v = constant(call(d, "constantTexture"))
    assert ch == 4 and np.array_equal(v, np.float32([0.25, 0.5, 0.75, 1.0]))
    ch, v = constant(call(r, "constantTexture"))
    assert ch == 1 and v[0] == np.float32(0.125)

    sym, pay = p["glass"]  # GlassData{reflection, transmission, roughnessX, roughnessY}
    assert sym == "glassInit" and len(pay) == 32
    h = struct.unpack("<4Q", pay)
    assert [constant(call(x, "constantTexture"))[0] for x in h] == [4, 4, 1, 1]
    assert np.allclose(constant(call(h[1], "constantTexture"))[1][:3], [0.6, 0.5, 0.4])

    sym, pay = p["checkered"]  # CheckBoardData{float scale, uint32 channel, float black[4], white[4]}
    cb = call(struct.unpack_from("<Q", pay)[0], "checkBoard")
    assert len(cb) == 40 and struct.unpack_from("<fI", cb) == (4.0, 4)
    assert np.allclose(np.frombuffer(cb, np.float32, 8, 8), [0.1, 0.2, 0.3, 1, 0.9, 0.8, 0.7, 1])

    sym, pay = p["area"]  # DiffuseAreaLightData{Spectrum radiance @0, Call sample @16, float area @24}
    assert sym == "diffuseInit" and len(pay) == 32
    assert struct.unpack_from("<3f", pay, 0) == (10.0, 20.0, 30.0)
    cdf = call(struct.unpack_from("<Q", pay, 16)[0], "planeSample")
    assert struct.unpack_from("<f", pay, 24)[0] == 2.5  # 0.5 x 2 + 1.5 x 1
    # CDFData{uint32 primitives @0, TraversalHandle @8, uint32 cdf @16, pdf @20, float inverseArea @24, uint32 size @28}
    assert len(cdf) == 32 and struct.unpack_from("<Q", cdf, 8)[0] == TRAVERSAL
    prim, c, pd = struct.unpack_from("<I", cdf, 0)[0], *struct.unpack_from("<II", cdf, 16)
    assert len({prim, c, pd}) == 3
    assert struct.unpack_from("<f", cdf, 24)[0] == np.float32(1 / 2.5) and struct.unpack_from("<I", cdf, 28)[0] == 2

    for name, init, vals in (("environment", "SEInit", (0.1, 0.2, 0.3)), ("point", "pointInit", (4.0, 5.0, 6.0))):
        sym, pay = p[name]  # {TraversalHandle, Spectrum}
        assert sym == init and len(pay) == 24 and struct.unpack_from("<Q", pay)[0] == TRAVERSAL
        assert np.allclose(struct.unpack_from("<3f", pay, 8), vals)

    sym, pay = p["sensor"]  # PCData{traversal, lookAt, size, upRef, focalLength, apertureRadius}
    assert sym == "rayGen" and len(pay) == 48 and struct.unpack_from("<Q", pay)[0] == TRAVERSAL
    assert np.allclose(struct.unpack_from("<3f2f3f2f", pay, 8), [1, 2, 3, 0.036, 0.024, 0, 1, 0, 0.05, 0.0125])

    sym, pay = p["integrator"]  # Data{uint32 maxDepth, uint32 minRR, 3 statistics handles}
    assert sym == "trace" and len(pay) == 32 and struct.unpack_from("<II", pay) == (12, 3)

    spp, dims, sobol = p["sampler"]  # SobolData{resolution, log2Resolution}; spp rounded up to 2^k
    assert (spp, dims) == (64, 1024) and struct.unpack("<II", sobol) == (2048, 11)

    sym, pay = p["lightsampler"]  # {uint32 count, bool delta[count]}
    assert sym == "uniformSelect" and pay == struct.pack("<I", 3) + bytes([0, 0, 1])

    sym, pay = p["driver"]  # RDData{Call filter} -> GaussianFilterData{negAlpha, sub}
    assert sym == "accumulate" and len(pay) == 8
    g = call(struct.unpack("<Q", pay)[0], "gaussian")
    assert struct.unpack("<ff", g) == (np.float32(-2.0), np.float32(np.exp(np.float32(-2.0))))
