"""The Rust binding (vistio_amd/rust/vistio-hip) against the C header.

There is no cargo here, so the crate is not compiled; what can break it
silently is a declaration that drifts from include/vistio_hip.h. This test
parses both: every C function has exactly one Rust declaration with the same
parameters (types mapped C -> Rust), every C struct has a #[repr(C)] Rust
struct with the same fields in the same order and types, and the enum
constants agree.
"""
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "vistio_hip.h")
CRATE = os.path.join(ROOT, "vistio_amd", "rust", "vistio-hip")
FFI = os.path.join(CRATE, "src", "ffi.rs")

_SCALAR = {"int32_t": "i32", "uint32_t": "u32", "int64_t": "i64", "uint64_t": "u64", "double": "f64",
           "float": "f32", "uint8_t": "u8", "char": "c_char", "vh_status": "VhStatus"}


def _camel(c_name: str) -> str:
    return "".join(p.capitalize() for p in c_name.split("_"))


def _rust_of(c_type: str, dims=None) -> str:
    """C type text ('const float*', 'vh_ctx**', 'double' with dims ['3'])."""
    t = c_type.replace("*", " * ").split()
    const = t[0] == "const"
    if const:
        t = t[1:]
    base, stars = t[0], t.count("*")
    r = _SCALAR.get(base) or (_camel(base) if base.startswith("vh_") else None)
    assert r, f"no Rust type for C type {c_type!r}"
    for k in range(stars):
        r = ("*const " if const and k == 0 else "*mut ") + r
    for d in reversed(dims or []):
        r = f"[{r}; {d}]"
    return r


def _strip_c_comments(s: str) -> str:
    s = re.sub(r"/\*.*?\*/", " ", s, flags=re.S)
    return re.sub(r"//[^\n]*", " ", s)


def _c_decl(text: str):
    """'const float* px' -> ('const float*', 'px', dims)."""
    m = re.match(r"^(.*?)(\w+)\s*((?:\[\w+\]\s*)*)$", text.strip())
    assert m, text
    ty, name, dims = m.group(1).strip(), m.group(2), re.findall(r"\[(\w+)\]", m.group(3))
    return ty, name, dims


def parse_header():
    src = _strip_c_comments(open(HEADER).read())
    funcs = {}
    for m in re.finditer(r"(vh_status|const\s+char\s*\*)\s*(vh_\w+)\s*\(([^)]*)\)\s*;", src):
        ret = "VhStatus" if m.group(1) == "vh_status" else "*const c_char"
        params = [_rust_of(*_c_decl(p)[::2]) for p in m.group(3).split(",") if p.strip()]
        funcs[m.group(2)] = (params, ret)
    structs = {}
    for m in re.finditer(r"typedef\s+struct\s*\{(.*?)\}\s*(\w+)\s*;", src, flags=re.S):
        fields = []
        for decl in m.group(1).split(";"):
            decl = decl.strip()
            if not decl:
                continue
            first, *rest = [d.strip() for d in decl.split(",")]
            ty, name, dims = _c_decl(first)
            fields.append((name, _rust_of(ty, dims)))
            base = ty.replace("*", "").strip()
            for extra in rest:  # 'double a, b' / 'float* x; float* y' style declarators
                stars = extra.count("*")
                _, nm, dm = _c_decl(extra.replace("*", ""))
                fields.append((nm, _rust_of(base + "*" * (stars + ty.count("*")), dm)))
        structs[_camel(m.group(2))] = fields
    consts = {k: int(v) for k, v in re.findall(r"\b(VH_[A-Z0-9_]+)\s*=\s*(\d+)", src)}
    consts.update({k: int(v) for k, v in re.findall(r"#define\s+(VH_[A-Z0-9_]+)\s+(\d+)", src)})
    return funcs, structs, consts


def parse_ffi():
    src = re.sub(r"//[^\n]*", " ", open(FFI).read())
    funcs = {}
    block = re.search(r'extern\s+"C"\s*\{(.*)\}', src, flags=re.S).group(1)
    for m in re.finditer(r"pub\s+fn\s+(vh_\w+)\s*\((.*?)\)\s*->\s*([^;]+?)\s*;", block, flags=re.S):
        params = []
        for p in m.group(2).split(","):
            if p.strip():
                params.append(" ".join(p.split(":", 1)[1].split()))
        funcs[m.group(1)] = (params, " ".join(m.group(3).split()))
    structs = {}
    for m in re.finditer(r"#\[repr\(C\)\]\s*(?:#\[derive\([^\]]*\)\]\s*)?pub\s+struct\s+(\w+)\s*\{(.*?)\}", src,
                         flags=re.S):
        fields = []
        for f in m.group(2).split(","):
            f = f.strip()
            if f and not f.startswith("_"):
                name, ty = f.split(":", 1)
                fields.append((name.replace("pub", "").strip(), " ".join(ty.split())))
        structs[m.group(1)] = fields
    consts = {k: int(v) for k, v in re.findall(r"pub\s+const\s+(VH_\w+)\s*:\s*\w+\s*=\s*(\d+)\s*;", src)}
    return funcs, structs, consts


def test_every_function_declared_alike():
    cf, _, _ = parse_header()
    rf, _, _ = parse_ffi()
    assert len(cf) >= 30
    assert sorted(cf) == sorted(rf), (set(cf) ^ set(rf))
    for name, (params, ret) in cf.items():
        assert rf[name] == (params, ret), f"{name}: header {params} -> {ret}, binding {rf[name]}"


def test_every_struct_laid_out_alike():
    _, cs, _ = parse_header()
    _, rs, _ = parse_ffi()
    assert set(cs) <= set(rs), set(cs) - set(rs)
    for name, fields in cs.items():
        assert rs[name] == fields, f"{name}:\n header  {fields}\n binding {rs[name]}"


def test_constants_agree():
    _, _, cc = parse_header()
    _, _, rc = parse_ffi()
    assert cc == rc, (cc, rc)


def test_parser_catches_a_drift():
    """A reordered field or a changed parameter type must fail the checks."""
    _, cs, _ = parse_header()
    f = list(cs["VhIpcParams"])
    f[0], f[1] = f[1], f[0]
    assert f != cs["VhIpcParams"]
    assert _rust_of("const double*") == "*const f64" and _rust_of("double*") == "*mut f64"
    assert _rust_of("vh_ctx**") == "*mut *mut VhCtx" and _rust_of("double", ["VH_MAX_BANDS"]) == "[f64; VH_MAX_BANDS]"


def test_crate_has_every_module_and_no_placeholders():
    src = os.path.join(CRATE, "src")
    mods = set(os.listdir(src))
    assert {"lib.rs", "ffi.rs", "ctx.rs", "solver.rs", "contact.rs", "backend.rs"} <= mods
    lib = open(os.path.join(src, "lib.rs")).read()
    for m in ("backend", "contact", "ctx", "ffi", "solver"):
        assert re.search(rf"\bmod {m};", lib), m
    for fn in mods:
        text = open(os.path.join(src, fn)).read()
        assert "todo!" not in text and "unimplemented!" not in text, fn
        assert text.count("{") == text.count("}"), f"unbalanced braces in {fn}"
        assert text.count("(") == text.count(")"), f"unbalanced parentheses in {fn}"
    for trait in ("impl SolverStrategy for HipPdSolver", "impl BroadPhase for HipBroadPhase",
                  "impl GpuBackend for HipBackend", "impl IpcCollisionHandler for HipIpcHandler"):
        assert any(trait in open(os.path.join(src, fn)).read() for fn in mods), trait
