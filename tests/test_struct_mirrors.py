"""The C ABI's structs (include/xpbd.h) against their mirrors: the Python
ctypes structures in bevy_xpbd_amd/api.py and the Rust #[repr(C)] structs in
INTEGRATION.md.  A field missing or out of order in a mirror shifts every
later field, so the names are compared in declaration order (no GPU)."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _c_structs():
    text = open(os.path.join(ROOT, "include", "xpbd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    out = {}
    for name, body in re.findall(r"typedef struct (\w+) \{(.*?)\} \1;", text, flags=re.S):
        fields = []
        for decl in body.split(";"):
            decl = " ".join(decl.split())
            if not decl:
                continue
            # "float dt, inv_dt" / "const uint32_t* ground" / "float gravity[4]"
            m = re.match(r"(?:const )?[\w]+\*? ?(.*)", decl)
            for part in m.group(1).split(","):
                fields.append(re.sub(r"[\s*]|\[.*\]", "", part))
        out[name] = fields
    return out


def _rust_structs():
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    out = {}
    # multi-line structs only (opaque handles are one-line "pub struct xpbd_mesh { .. }")
    for name, body in re.findall(r"pub struct (xpbd_\w+) \{(?![^\n]*\})[^\n]*\n(.*?)\n\}", text, flags=re.S):
        body = re.sub(r"//[^\n]*", "", body)
        out[name] = re.findall(r"pub (\w+):", body)
    return out


MIRRORS = [("xpbd_sheet_desc", "SheetDesc"), ("xpbd_mesh_desc", "MeshDesc"), ("xpbd_sim_config", "ClothSimConfig"),
           ("xpbd_sim_uniforms", "ClothSimUniforms")]


def test_header_parses():
    c = _c_structs()
    for cname, _ in MIRRORS:
        assert cname in c and len(c[cname]) >= 4, cname
    assert c["xpbd_sim_uniforms"][:4] == ["dt", "inv_dt", "num_particles", "num_tris"]


@pytest.mark.parametrize("cname,pyname", MIRRORS)
def test_python_mirror_fields_in_order(X, cname, pyname):
    from bevy_xpbd_amd import api
    py = [f[0] for f in getattr(api, pyname)._fields_]
    assert py == _c_structs()[cname]


@pytest.mark.parametrize("cname", ["xpbd_mesh_desc", "xpbd_sim_config", "xpbd_sim_uniforms"])
def test_rust_mirror_fields_in_order(cname):
    rust = _rust_structs()
    assert cname in rust, sorted(rust)
    assert rust[cname] == _c_structs()[cname]
