"""The C ABI library loads and exports every symbol include/aemh.h declares
(no compute calls: these run without a GPU)."""
import ctypes
import os
import re
import subprocess

import numpy as np

import aemh

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    src = open(os.path.join(ROOT, "include", "aemh.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(aemh_[a-z0-9_]+)\s*\(", src)))


def test_header_symbols_listed():
    assert header_symbols() == sorted(aemh.EXPORTS)


def test_library_exports_all_symbols():
    lib = ctypes.CDLL(aemh.LIB_PATH)
    for name in header_symbols():
        assert hasattr(lib, name), name


def test_version_and_defaults():
    lib = aemh.load_library()
    assert b"gfx950" in lib.aemh_version()
    cfg = aemh.default_config()
    # MpmConfig::default (ae_particle/src/mpm_solver.rs:57-69)
    assert cfg.substeps == 1
    assert tuple(cfg.gravity) == (0.0, np.float32(-9.81), 0.0)
    assert abs(cfg.damping - 0.995) < 1e-7
    assert cfg.accum == aemh.ACCUM_F32 and cfg.clamp_mode == aemh.CLAMP_AUTO


def test_struct_layouts_match_header(tmp_path):
    """ctypes mirrors of aemh_config / aemh_scene have the C compiler's layout."""
    src = tmp_path / "layout.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "aemh.h"\n'
        'int main(void){printf("%zu %zu %zu %zu %zu\\n", sizeof(aemh_config), sizeof(aemh_scene),'
        ' offsetof(aemh_config, accum), offsetof(aemh_scene, dt), offsetof(aemh_scene, clamp_mode));'
        'return 0;}\n')
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = [int(v) for v in subprocess.run([str(exe)], check=True, capture_output=True,
                                          text=True).stdout.split()]
    assert out == [ctypes.sizeof(aemh.Config), ctypes.sizeof(aemh.Scene),
                   aemh.Config.accum.offset, aemh.Scene.dt.offset, aemh.Scene.clamp_mode.offset]


def test_scene_generator_is_seeded_and_bounded():
    a = aemh.uniform_box(42, 1000, (0.3, 0.3, 0.3), (0.6, 0.6, 0.6))
    b = aemh.uniform_box(42, 1000, (0.3, 0.3, 0.3), (0.6, 0.6, 0.6))
    c = aemh.uniform_box(43, 1000, (0.3, 0.3, 0.3), (0.6, 0.6, 0.6))
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert a.min() >= 0.3 and a.max() < 0.6
    # uniformity: each axis mean ~0.45 within 5 sigma
    assert np.all(np.abs(a.mean(0) - 0.45) < 5 * 0.3 / np.sqrt(12 * 1000))


def test_create_without_gpu_fails_cleanly():
    import torch
    if torch.cuda.device_count() > 0:
        return
    try:
        aemh.Engine()
    except aemh.AemhError as e:
        assert e.status in (aemh.ERR_NO_DEVICE, aemh.ERR_HIP)
    else:
        raise AssertionError("Engine() must fail without a gfx950 device")


def _param_counts(src, pattern):
    """name -> number of parameters of every declaration matching pattern."""
    out = {}
    for m in re.finditer(pattern, src):
        depth, i = 1, m.end()
        while depth:
            depth += {"(": 1, ")": -1}.get(src[i], 0)
            i += 1
        args = src[m.end():i - 1].strip()
        out[m.group(1)] = 0 if args in ("", "void") else args.count(",") + 1
    return out


def test_rust_binding_declares_every_entry_point():
    """ae_mpm_hip (the Rust crate a maintainer adds to the workspace; there is
    no cargo here) declares every function of include/aemh.h with the same
    number of parameters, and its aemh_config / aemh_scene mirrors have the
    header's fields in order."""
    hdr = re.sub(r"/\*.*?\*/", "", open(os.path.join(ROOT, "include", "aemh.h")).read(), flags=re.S)
    rs = open(os.path.join(ROOT, "ae-engine_amd", "ae_mpm_hip", "src", "lib.rs")).read()
    want = _param_counts(hdr, r"\b(aemh_[a-z0-9_]+)\s*\(")
    got = _param_counts(rs, r"pub fn (aemh_[a-z0-9_]+)\s*\(")
    assert got == want
    for c_name, rs_name in (("aemh_config", "AemhConfig"), ("aemh_scene", "AemhScene")):
        c_body = re.search(r"typedef struct \{([^}]*)\} " + c_name + ";", hdr).group(1)
        c_fields = []
        for decl in c_body.split(";")[:-1]:
            names = decl.split(None, 1)[1]  # drop the type
            c_fields += [re.match(r"\s*([a-z_0-9]+)", d).group(1) for d in names.split(",")]
        rs_body = re.search(r"pub struct " + rs_name + r" \{(.*?)\}", rs, re.S).group(1)
        rs_fields = re.findall(r"pub ([a-z_0-9]+):", rs_body)
        assert c_fields == rs_fields, (c_name, c_fields, rs_fields)
