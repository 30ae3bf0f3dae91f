"""The C-ABI library loads and exports every entry point include/vkhr_hip.h
declares; struct layouts match the header.  No compute calls (CPU only)."""
import ctypes as C
import re

import pytest

from vkhr_amd import capi


def test_every_declared_symbol_is_exported():
    declared = capi.declared_symbols()
    assert "vkhr_voxelize_segments" in declared
    missing = [s for s in declared if not hasattr(capi.lib, s)]
    assert not missing, f"declared but not exported: {missing}"


def test_every_declared_symbol_has_a_binding():
    declared = set(capi.declared_symbols()) - {"vkhr_last_error"}
    bound = set(capi.SIGNATURES) | set(capi.SIZE_FUNCTIONS)
    assert declared == bound, declared ^ bound


def test_abi_version_and_error_string():
    assert capi.lib.vkhr_abi_version() == 3
    assert capi.lib.vkhr_march_record_bytes(16, 4) == 64 * 64
    assert isinstance(capi.lib.vkhr_last_error(), bytes)


def _header_struct_fields(name):
    text = capi.HEADER.read_text()
    m = re.search(r"typedef struct %s \{(.*?)\} %s;" % (name, name), text, re.S)
    assert m, name
    body = re.sub(r"/\*.*?\*/", "", m.group(1), flags=re.S)
    fields = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl:
            continue
        words = decl.replace(",", " , ").replace("*", " ").replace("const ", " ").split()
        names = [w for w in words[1:] if w != ","]
        for n in names:
            base = n.split("[")[0]
            count = int(n.split("[")[1].rstrip("]")) if "[" in n else 1
            fields.append((base, count))
    return fields


def test_struct_layouts_match_header():
    pairs = [("vkhr_volume", capi.Volume), ("vkhr_aabb", capi.AABB), ("vkhr_camera", capi.Camera),
             ("vkhr_light", capi.Light), ("vkhr_strand_params", capi.StrandParams),
             ("vkhr_render_params", capi.RenderParams), ("vkhr_ppll", capi.Ppll),
             ("vkhr_frag_capture", capi.FragCapture), ("vkhr_diff_params", capi.DiffParams),
             ("vkhr_batch_style", capi.BatchStyle), ("vkhr_tuning", capi.Tuning),
             ("vkhr_adjoint_style", capi.AdjointStyle), ("vkhr_adjoint_render", capi.AdjointRender)]
    for cname, cls in pairs:
        hdr = _header_struct_fields(cname)
        names = [f[0] for f in cls._fields_]
        assert [h[0] for h in hdr] == names, cname
    assert C.sizeof(capi.AABB) == 32
    assert C.sizeof(capi.Camera) == 4 * (16 + 16 + 3 + 1 + 2 + 2)
    assert C.sizeof(capi.Light) == 4 * (4 + 4 + 16 + 3 + 1 + 1 + 3)
    assert C.sizeof(capi.StrandParams) == 32 + 4 * 10
    assert C.sizeof(capi.RenderParams) == 4 * 19


def test_init_without_gpu_reports_an_error_code():
    import torch
    if torch.cuda.is_available():
        return
    rc = capi.lib.vkhr_init()
    assert rc < 0
    assert capi.lib.vkhr_last_error()


def test_misplaced_arguments_are_refused_before_the_call():
    """A pointer-sized value in a 32-bit slot or a small integer in a pointer
    slot raises on the host (no device work is enqueued)."""
    from vkhr_amd import capi
    ptr = 0x7F0000001000
    with pytest.raises(ValueError, match="not a pointer"):
        capi.call("vkhr_bvh_build", ptr, 4000, ptr, ptr, 8, None, None)
    with pytest.raises(ValueError, match="does not fit 32 bits"):
        capi.call("vkhr_bvh_build", ptr, ptr, ptr, ptr, 8, None, None)
