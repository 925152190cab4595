"""The Rust binding a maintainer adds (integration/rust/rc-gi-sys) against
the C header it binds: every entry point of include/rc_gi.h is declared with
the same parameter and return types, every POD struct has the same fields
in the same order with the same types, and the constants agree.  cargo is
not in this image, so this is the check that the binding cannot drift from
the ABI; the stage module (integration/rust/gi_stage.rs) only calls through
these declarations."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "rc_gi.h")
SYS = os.path.join(ROOT, "integration", "rust", "rc-gi-sys", "src", "lib.rs")
STAGE = os.path.join(ROOT, "integration", "rust", "gi_stage.rs")

SCALAR = {"uint32_t": "u32", "int32_t": "i32", "int": "c_int", "uint64_t": "u64", "int64_t": "i64",
          "uint8_t": "u8", "float": "f32", "double": "f64", "char": "c_char", "void": "c_void",
          "rc_status": "rc_status", "size_t": "usize"}


def _header():
    return re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)


def c_type_to_rust(t: str) -> str:
    """'const rc_config *' -> '*const rc_config', 'rc_ctx **' -> '*mut *mut rc_ctx'."""
    t = " ".join(t.replace("*", " * ").split())
    stars = t.count("*")
    base = t.replace("*", "").split()
    const = base[0] == "const"
    if const:
        base = base[1:]
    name = SCALAR.get(base[0], base[0])
    if stars == 0:
        return name
    out = ("*const " if const else "*mut ") + name
    for _ in range(stars - 1):
        out = "*mut " + out
    return out


def c_functions():
    out = {}
    for ret, name, args in re.findall(r"^((?:const\s+)?\w+(?:\s*\*)*)\s*(rc_\w+)\s*\(([^;]*?)\)\s*;", _header(), re.M):
        params = []
        for a in " ".join(args.split()).split(","):
            a = a.strip()
            if a in ("", "void"):
                continue
            m = re.match(r"(.*?)(\w+)$", a)
            params.append(c_type_to_rust(m.group(1)))
        r = c_type_to_rust(ret)
        out[name] = (params, None if r == "c_void" else r)
    return out


def rust_functions(src=None):
    src = open(SYS).read() if src is None else src
    block = src[src.index('extern "C" {'):]
    block = re.sub(r"//[^\n]*", "", block)
    out = {}
    for name, args, ret in re.findall(r"pub fn (rc_\w+)\s*\((.*?)\)\s*(?:->\s*([^;]+))?;", block, re.S):
        params = []
        for a in " ".join(args.split()).split(","):
            a = a.strip()
            if a:
                params.append(a.split(":", 1)[1].strip())
        out[name] = (params, ret.strip() if ret else None)
    return out


def test_every_entry_point_is_bound_with_the_same_types():
    c = c_functions()
    r = rust_functions()
    assert len(c) >= 60
    assert set(c) == set(r), (sorted(set(c) - set(r)), sorted(set(r) - set(c)))
    for name in c:
        assert r[name] == c[name], (name, c[name], r[name])


def _c_structs():
    out = {}
    for body, name in re.findall(r"typedef struct\s*\{(.*?)\}\s*(\w+);", _header(), re.S):
        fields = []
        for decl in body.split(";"):
            decl = " ".join(decl.split())
            if not decl:
                continue
            m = re.match(r"((?:const\s+)?\w+\s*\**)\s*(.*)", decl)
            base, names = m.group(1), m.group(2)
            for n in names.split(","):
                n = n.strip()
                stars = n.count("*")
                n = n.replace("*", "").strip()
                arr = re.match(r"(\w+)\[(\w+)\]", n)
                ty = c_type_to_rust(base + "*" * stars)
                if arr:
                    fields.append((arr.group(1), f"[{ty}; {arr.group(2)}]"))
                else:
                    fields.append((n, ty))
        out[name] = fields
    return out


def _rust_structs():
    src = open(SYS).read()
    out = {}
    for name, body in re.findall(r"pub struct (rc_\w+)\s*\{(.*?)\n\}", src, re.S):
        fields = re.findall(r"pub (\w+): ([^,]+),", body)
        if fields:
            out[name] = [(n, " ".join(t.split())) for n, t in fields]
    return out


def test_every_struct_has_the_same_fields():
    c = _c_structs()
    r = _rust_structs()
    assert set(c) == set(r), (sorted(set(c) - set(r)), sorted(set(r) - set(c)))
    for name in c:
        cf = [(n, t.replace("[c_char; 16]", "[c_char; 16]")) for n, t in c[name]]
        rf = [(n, t.replace("RC_NUM_STAGES", "RC_NUM_STAGES").replace("RC_MAX_LEVELS", "RC_MAX_LEVELS")) for n, t in r[name]]
        assert [n for n, _ in cf] == [n for n, _ in rf], name
        assert [t for _, t in cf] == [t for _, t in rf], (name, cf, rf)


def test_constants_agree():
    h = _header()
    src = open(SYS).read()
    for name in ("RC_ABI_VERSION", "RC_MAX_LEVELS", "RC_MAX_DIM", "RC_MAX_BOXES", "RC_FLAG_PROFILE",
                 "RC_FLAG_KEEP_DEBUG", "RC_FLAG_GRAPH", "RC_DEPTH_LINEAR", "RC_DEPTH_NDC", "RC_NUM_TILE_SHAPES",
                 "RC_MLP_INPUTS", "RC_MLP_OUTPUTS", "RC_MLP_PARAMS", "RC_TRAIN_MAX_EPOCHS", "RC_BATCH_WORKLIST",
                 "RC_BATCH_STREAMS", "RC_STAGE_SEED", "RC_STAGE_JFA", "RC_STAGE_JFA_TAIL", "RC_STAGE_LEVEL0"):
        cv = re.search(rf"#define {name}\s+([0-9a-fA-Fx]+)u?\b", h).group(1)
        rv = re.search(rf"pub const {name}: \w+ = ([0-9a-fA-Fx]+);", src).group(1)
        assert int(cv, 0) == int(rv, 0), (name, cv, rv)
    # the stage offsets are derived the same way
    for name in ("RC_STAGE_MARCH", "RC_STAGE_NEAR_SDF", "RC_STAGE_MIP", "RC_STAGE_CONE", "RC_STAGE_SURROGATE",
                 "RC_STAGE_FALLBACK", "RC_NUM_STAGES"):
        cv = re.search(rf"#define {name}\s+\(([^)]*)\)", h).group(1)
        rv = re.search(rf"pub const {name}: usize = ([^;]+);", src).group(1)
        assert " ".join(cv.split()) == " ".join(rv.split()), (name, cv, rv)


def test_stage_module_uses_only_bound_entry_points():
    """gi_stage.rs implements the reference trait's three methods and calls
    the library only through rc-gi-sys declarations."""
    src = open(STAGE).read()
    bound = set(rust_functions())
    called = set(re.findall(r"\b(rc_\w+)\s*\(", src))
    assert called and called <= bound, sorted(called - bound)
    for m in ("fn render(&self, _state: &mut AppState, _view: &wgpu::TextureView, _encoder: &mut wgpu::CommandEncoder)",
              "fn resize(&mut self, device: &wgpu::Device, config: &wgpu::SurfaceConfiguration)",
              "fn update(&mut self, state: &AppState, queue: &wgpu::Queue)"):
        assert m in src, m
    assert "depth_mode: RC_DEPTH_NDC" in src and "z_near: 0.1" in src and "z_far: 100.0" in src


def test_the_check_sees_a_planted_mistake():
    """A parameter of the wrong width, a const-ness flip and a dropped
    declaration in a copy of the binding each fail the comparison."""
    c = c_functions()
    src = open(SYS).read()
    for bad in (src.replace("pub fn rc_resize(ctx: *mut rc_ctx, width: u32", "pub fn rc_resize(ctx: *mut rc_ctx, width: u64"),
                src.replace("pub fn rc_update_host(ctx: *mut rc_ctx, host: *const rc_gbuffer)",
                            "pub fn rc_update_host(ctx: *mut rc_ctx, host: *mut rc_gbuffer)"),
                src.replace("    pub fn rc_batch_run(batch: *mut rc_batch, stream: *mut c_void) -> rc_status;\n", "")):
        assert bad != src
        r = rust_functions(bad)
        assert set(c) != set(r) or any(r[n] != c[n] for n in c)
