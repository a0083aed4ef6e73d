"""The Rust FFI declarations written out in INTEGRATION.md against the C headers (include/*.h): no
Rust toolchain exists in this image, so the crate cannot be compiled here, but every `#[repr(C)]`
struct must list the header's fields in the header's order with the same types, and every
function declared in an `extern "C"` block must exist in a header with the same number of
parameters. A struct or signature that drifts from the header would corrupt every call."""
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

RUST_TO_C = {"i32": "int32_t", "u32": "uint32_t", "i64": "int64_t", "u64": "uint64_t", "f32": "float", "f64": "double",
             "u8": "uint8_t", "i8": "int8_t", "u16": "uint16_t", "c_int": "int", "usize": "size_t"}


def _strip_c_comments(text):
    return re.sub(r"/\*.*?\*/", " ", re.sub(r"//[^\n]*", " ", text), flags=re.S)


def header_structs():
    """{name: [(field, C type)]} from every `typedef struct name { ... } name;` of include/*.h."""
    out = {}
    for fn in sorted(os.listdir(os.path.join(ROOT, "include"))):
        text = _strip_c_comments(open(os.path.join(ROOT, "include", fn)).read())
        for m in re.finditer(r"typedef\s+struct\s+(\w+)\s*\{(.*?)\}\s*(\w+)\s*;", text, re.S):
            fields = []
            for decl in m.group(2).split(";"):
                decl = " ".join(decl.split())
                if not decl:
                    continue
                dm = re.match(r"(.*?)(\w+)\s*(\[\s*(\w+)\s*\])?$", decl)
                ctype = dm.group(1).replace(" *", "*").strip()
                if dm.group(4):
                    ctype = f"{ctype}[{dm.group(4)}]"
                fields.append((dm.group(2), ctype))
            out[m.group(1)] = fields
    return out


def header_functions():
    """{name: number of parameters} of every function declared in include/*.h."""
    out = {}
    for fn in sorted(os.listdir(os.path.join(ROOT, "include"))):
        text = _strip_c_comments(open(os.path.join(ROOT, "include", fn)).read())
        for m in re.finditer(r"\b(c3d_\w+)\s*\(([^;{]*?)\)\s*;", text, re.S):
            args = m.group(2).strip()
            out[m.group(1)] = 0 if args in ("", "void") else len(args.split(","))
    return out


def rust_blocks():
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    return re.findall(r"```rust\n(.*?)```", text, re.S)


def _rust_type(t, consts):
    t = " ".join(t.split())
    m = re.match(r"\[(.+);\s*(\w+)\]$", t)
    if m:
        n = consts.get(m.group(2), m.group(2))
        return f"{_rust_type(m.group(1), consts)}[{n}]"
    m = re.match(r"\*(const|mut) (.+)$", t)
    if m:
        inner = _rust_type(m.group(2), consts)
        return ("const " if m.group(1) == "const" else "") + inner + "*"
    return RUST_TO_C.get(t, t)


def rust_structs():
    """{name: [(field, C-spelled type)]} of every `#[repr(C)]` struct with named fields."""
    out = {}
    for block in rust_blocks():
        consts = dict(re.findall(r"pub const (\w+): \w+ = (\w+);", block))
        for m in re.finditer(r"#\[repr\(C\)\](?:\s*#\[derive\([^)]*\)\])?\s*pub struct (\w+)\s*\{(.*?)\}", block, re.S):
            body = m.group(2).strip()
            if body.startswith("_p"):
                continue  # opaque handle
            fields = []
            for f in re.split(r",\s*(?=pub )", body):
                f = f.strip().rstrip(",")
                fm = re.match(r"pub (\w+):\s*(.+)$", f, re.S)
                fields.append((fm.group(1), _rust_type(fm.group(2), consts)))
            out[m.group(1)] = fields
    return out


def rust_functions():
    """{name: number of parameters} of every `fn` in the `extern "C"` blocks."""
    out = {}
    for block in rust_blocks():
        for ext in re.finditer(r'extern "C" \{(.*?)\n\}', block, re.S):
            for m in re.finditer(r"pub fn (\w+)\((.*?)\)", ext.group(1), re.S):
                args = [a for a in m.group(2).split(",") if a.strip()]
                out[m.group(1)] = len(args)
    return out


def _normalise_c(t):
    """Layout-relevant spelling: qualifiers dropped (const does not change the ABI), spaces removed."""
    return re.sub(r"\bconst\b", "", t).replace(" ", "")


def test_rust_structs_match_the_headers():
    hs = header_structs()
    rs = rust_structs()
    assert sorted(rs) == sorted(hs), f"only in C: {sorted(set(hs) - set(rs))}, only in Rust: {sorted(set(rs) - set(hs))}"
    for name, fields in rs.items():
        assert name in hs, f"{name} is not a header struct"
        assert [f for f, _ in fields] == [f for f, _ in hs[name]], f"{name}: field names or order differ"
        for (f, rt), (_, ct) in zip(fields, hs[name]):
            assert _normalise_c(rt) == _normalise_c(ct), f"{name}.{f}: Rust {rt} vs C {ct}"


def test_rust_functions_exist_with_the_header_arity():
    hf = header_functions()
    rf = rust_functions()
    assert sorted(rf) == sorted(hf), f"declared only in C: {sorted(set(hf) - set(rf))}, only in Rust: {sorted(set(rf) - set(hf))}"
    for name, n in rf.items():
        assert name in hf, f"{name} is declared in INTEGRATION.md but not in include/*.h"
        assert n == hf[name], f"{name}: {n} parameters in Rust, {hf[name]} in C"


def test_a_drifted_struct_is_caught():
    """The comparison itself: swapping two fields or changing a width fails it."""
    hs = header_structs()
    f = list(hs["c3d_run_params"])
    assert f[:2] == [("level", "uint32_t"), ("geometric", "int32_t")]
    swapped = [f[1], f[0]] + f[2:]
    assert [x for x, _ in swapped] != [x for x, _ in hs["c3d_run_params"]]
    assert _normalise_c(_rust_type("[f32; 9]", {})) == "float[9]"
    assert _normalise_c(_rust_type("*const *const f32", {})) == _normalise_c("const float* const*") == "float**"
    assert _normalise_c(_rust_type("*const f32", {})) != _normalise_c("const float* const*")
    assert _rust_type("*mut f32", {}) == "float*" and _rust_type("u64", {}) != "uint32_t"
