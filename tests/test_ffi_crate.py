"""The Rust FFI crate (jacquard_amd/ffi/jacquard-hip-sys) declares every
function of include/jacquard_hip.h, and only those, with the same
parameter names, count and C-to-Rust types.  No Rust toolchain is in this
image, so the crate is checked against the header as text; the header
itself is checked against the library's exports in
test_capi_vcd_events.py, and the same symbols are called from compiled C
in test_abi_c.py."""
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CRATE = os.path.join(ROOT, "jacquard_amd", "ffi", "jacquard-hip-sys")

C_TO_RUST = {"usize": "usize", "u32": "u32", "i32": "i32", "u8": "u8", "int": "c_int",
             "float": "f32", "void": "c_void", "size_t": "usize", "char": "c_char"}


def _c_type(t):
    t = " ".join(t.split())
    const = t.startswith("const ")
    base = t[6:] if const else t
    stars = base.count("*")
    r = C_TO_RUST[base.replace("*", "").strip()]
    for _ in range(stars):
        r = ("*const " if const else "*mut ") + r
    return r


def header_decls():
    text = open(os.path.join(ROOT, "include", "jacquard_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    out = {}
    for m in re.finditer(r"([A-Za-z_][\w \*]*?)\b([a-z_][a-z0-9_]*)\s*\(([^;{]*?)\)\s*;", text):
        ret, name, params = m.group(1).strip(), m.group(2), " ".join(m.group(3).split())
        ps = []
        if params != "void":
            for p in params.split(","):
                mm = re.match(r"(.*?)([A-Za-z_]\w*)$", p.strip())
                ps.append((mm.group(2), _c_type(mm.group(1))))
        out[name] = (ps, None if ret == "void" else _c_type(ret))
    return out


def crate_decls():
    text = open(os.path.join(CRATE, "src", "lib.rs")).read()
    block = text[text.index('extern "C" {'):]
    block = block[:block.index("\n}\n")]
    block = re.sub(r"///[^\n]*", "", block)
    out = {}
    for m in re.finditer(r"pub fn (\w+)\(([^)]*)\)\s*(->\s*([^;]+))?;", block):
        ps = []
        for p in m.group(2).split(","):
            p = p.strip()
            if p:
                n, t = p.split(":", 1)
                ps.append((n.strip().rstrip("_"), " ".join(t.split())))
        out[m.group(1)] = (ps, m.group(4).strip() if m.group(4) else None)
    return out


def test_crate_declares_exactly_the_header():
    h, r = header_decls(), crate_decls()
    assert "simulate_v1_noninteractive_simple_scan_hip" in h
    assert set(h) == set(r), (set(h) ^ set(r))
    for name, (ps, ret) in h.items():
        rps, rret = r[name]
        assert [n for n, _ in rps] == [n for n, _ in ps], name
        assert [t for _, t in rps] == [t for _, t in ps], name
        assert rret == ret, name


def test_crate_links_the_library():
    build = open(os.path.join(CRATE, "build.rs")).read()
    assert "rustc-link-lib=dylib=jacquard_hip" in build
    assert "rustc-link-lib=dylib=amdhip64" in build
    manifest = open(os.path.join(CRATE, "Cargo.toml")).read()
    assert 'links = "jacquard_hip"' in manifest
