"""The C ABI: struct layouts, exported symbols, and the host loader's error behaviour. CPU only.

No test here makes a HIP call: the plugin library is loaded and its pure version entry points
are called, nothing else.
"""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import prender_amd as pa

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")


def declared(header, macro):
    text = open(os.path.join(INCLUDE, header)).read()
    return re.findall(macro + r"\s+[\w:\s\*]+?\b(\w+)\s*\(", text)


def test_every_declared_plugin_symbol_is_exported():
    names = declared("prender_gpu.h", "PRGPU_API")
    # The three v4 entry points of the reference plugin (src/integrators/gpu_path.cpp:69-80) come first.
    assert names[:3] == ["prgpu_list_devices", "prgpu_render", "prgpu_api_version"]
    lib = C.CDLL(pa.PLUGIN_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert len(names) >= 12


def test_every_declared_host_symbol_is_exported():
    names = declared("prender_host.h", "PH_API")
    lib = C.CDLL(pa.HOST_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_only_declared_symbols_are_exported():
    for path, header, macro in ((pa.PLUGIN_PATH, "prender_gpu.h", "PRGPU_API"),
                                (pa.HOST_PATH, "prender_host.h", "PH_API")):
        out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True, check=True).stdout
        text_syms = {ln.split()[-1] for ln in out.splitlines() if ln.split()[1:2] == ["T"]}
        assert text_syms == set(declared(header, macro)), (path, text_syms ^ set(declared(header, macro)))


def test_versions_without_gpu():
    lib = C.CDLL(pa.PLUGIN_PATH)
    assert lib.prgpu_api_version() == pa.API_VERSION == 4
    assert lib.prgpu_api_ext_version() == pa.EXT_VERSION


def test_sheen_table_matches_oracle_without_gpu():
    """The plugin's built-in sheen albedo table is the CPU path tracer's, bit for bit."""
    import oracle as orc
    lib = C.CDLL(pa.PLUGIN_PATH)
    t = np.zeros(16 * 32, np.float32)
    assert lib.prgpu_sheen_table(C.c_void_p(t.ctypes.data)) == 512
    assert np.array_equal(t.view(np.uint32), orc.sheen_table().view(np.uint32))


def test_struct_sizes_match_header():
    text = open(os.path.join(INCLUDE, "prender_gpu.h")).read()
    sizes = dict(re.findall(r"static_assert\(sizeof\(prgpu::(\w+)\) == (\d+)", text))
    py = {"Primitive": pa.PRIM_DTYPE.itemsize, "Material": pa.MATERIAL_DTYPE.itemsize,
          "SpectrumEx": pa.SPECTRUM_DTYPE.itemsize, "LightEx": pa.LIGHT_DTYPE.itemsize,
          "MediumEx": pa.MEDIUM_DTYPE.itemsize, "SampleRecordEx": pa.SAMPLE_DTYPE.itemsize,
          "DeviceInfo": C.sizeof(pa.DeviceInfo), "ProfileStatsEx": C.sizeof(pa.ProfileStatsEx)}
    for name, size in py.items():
        assert int(sizes[name]) == size, name
    assert pa.RAY_DTYPE.itemsize == 28 and pa.HIT_DTYPE.itemsize == 16
    assert C.sizeof(pa.RenderParamsEx) == 40
    assert C.sizeof(pa.SceneDescEx) == 208


def test_loader_errors(tmp_path, host):
    err = C.create_string_buffer(512)
    assert not host.lib.ph_plugin_open(str(tmp_path / "missing.so").encode(), err, 512)
    assert err.value, "no message for a missing plugin"
    # A library without the plugin's symbols, or reporting another API version, is refused with the
    # reference loader's one status message (src/integrators/gpu_path.cpp:58-64).
    assert not host.lib.ph_plugin_open(pa.HOST_PATH.encode(), err, 512)
    assert b"incompatible interface" in err.value
    src = tmp_path / "old.c"
    src.write_text("int prgpu_api_version(void){return 3;}\n"
                   "int prgpu_list_devices(void*o,int m){return 0;}\n"
                   "int prgpu_render(int d,const void*s,const void*p,float*x,char*e,int l){return 1;}\n")
    so = tmp_path / "libold.so"
    subprocess.run(["gcc", "-shared", "-fPIC", "-o", str(so), str(src)], check=True)
    assert not host.lib.ph_plugin_open(str(so).encode(), err, 512)
    assert b"incompatible interface" in err.value


def test_v4_plugin_loads_without_extension(tmp_path, host):
    """A plugin with only the three v4 symbols opens, and reports no extension."""
    src = tmp_path / "v4.c"
    src.write_text("int prgpu_api_version(void){return 4;}\n"
                   "int prgpu_list_devices(void*o,int m){return 0;}\n"
                   "int prgpu_render(int d,const void*s,const void*p,float*x,char*e,int l){return 1;}\n")
    so = tmp_path / "libv4.so"
    subprocess.run(["gcc", "-shared", "-fPIC", "-o", str(so), str(src)], check=True)
    err = C.create_string_buffer(512)
    h = host.lib.ph_plugin_open(str(so).encode(), err, 512)
    assert h, err.value
    assert host.lib.ph_plugin_has_ext(h) == 0
    devs = (pa.DeviceInfo * 4)()
    assert host.lib.ph_plugin_list_devices(h, devs, 4) == 0
    host.lib.ph_plugin_close(h)


def test_render_params_defaults():
    p = pa.render_params()
    assert p.spp == 1 and p.seed == 1 and p.russianRoulette == 1


def test_sample_record_layout():
    rec = np.zeros(2, pa.SAMPLE_DTYPE)
    assert rec["L"].shape == (2, 4) and rec["pFilm"].shape == (2, 2)
