"""The Rust binding (crates/tinytrace_amd_sys/src/lib.rs, not compiled here:
no Rust toolchain in this image) against the C header it binds.

Both files are parsed into canonical types.  Every struct must have the same
fields in the same order with the same types, every function the same
parameter and return types, every constant the same value; and the offsets
and sizes that Rust's #[repr(C)] rules give for the Rust fields must equal
gcc's offsetof/sizeof for the C structs (a compiled probe of the header).
A binding that drops, reorders or retypes a field fails here.
"""
import os
import re
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "tinytrace.h")
CRATE = os.path.join(ROOT, "crates", "tinytrace_amd_sys", "src", "lib.rs")

C_SCALARS = {"uint8_t": "u8", "uint16_t": "u16", "int16_t": "i16", "uint32_t": "u32", "int32_t": "i32",
             "uint64_t": "u64", "float": "f32", "double": "f64", "size_t": "usize", "int": "i32",
             "char": "c_char", "tt_status": "tt_status"}
RUST_SCALARS = {"u8": "u8", "u16": "u16", "i16": "i16", "u32": "u32", "i32": "i32", "u64": "u64", "f32": "f32",
                "f64": "f64", "usize": "usize", "c_int": "i32", "c_char": "c_char", "tt_status": "tt_status"}
SIZES = {"u8": 1, "u16": 2, "i16": 2, "u32": 4, "i32": 4, "u64": 8, "f32": 4, "f64": 8, "usize": 8,
         "c_char": 1, "tt_status": 4}


def strip_c_comments(text):
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    return re.sub(r"//[^\n]*", " ", text)


def c_type(decl_type, dims=()):
    """Canonical type of a C declaration: base type text plus array dims."""
    t = decl_type.replace("struct ", "").strip()
    stars = t.count("*")
    base = t.replace("*", " ").split()
    const_levels = []
    # "const T*" -> pointer to const T; "T**" -> pointer to pointer to mut T.
    is_const = base[0] == "const"
    name = base[-1]
    canon = C_SCALARS.get(name, name)
    for level in range(stars):
        canon = ("ptr", "const" if (is_const and level == 0) else "mut", canon)
    for n in reversed(dims):
        canon = ("arr", canon, n)
    return canon


def parse_c_structs(text):
    structs = {}
    for m in re.finditer(r"typedef struct (\w+) \{(.*?)\} \1;", text, re.S):
        fields = []
        for decl in m.group(2).split(";"):
            decl = decl.strip()
            if not decl:
                continue
            dm = re.match(r"((?:const\s+)?(?:struct\s+)?\w+\s*\**)\s*(.*)$", decl)
            base, names = dm.group(1), dm.group(2)
            for item in names.split(","):
                item = item.strip()
                nm = re.match(r"(\**)\s*(\w+)((?:\[\d+\])*)$", item)
                assert nm, item
                dims = tuple(int(d) for d in re.findall(r"\[(\d+)\]", nm.group(3)))
                fields.append((nm.group(2), c_type(base + nm.group(1), dims)))
        structs[m.group(1)] = fields
    return structs


def rust_type(text):
    text = text.strip()
    if text.startswith("*const ") or text.startswith("*mut "):
        kind, rest = text[1:].split(" ", 1)
        return ("ptr", kind, rust_type(rest))
    if text.startswith("["):
        inner, n = text[1:-1].rsplit(";", 1)
        return ("arr", rust_type(inner), int(n))
    return RUST_SCALARS.get(text, text)


def parse_rust_structs(text):
    structs = {}
    for m in re.finditer(r"#\[repr\(C\)\]\n(?:#\[derive\([^)]*\)\]\n)?pub struct (\w+) \{\n(.*?)\n\}", text, re.S):
        body = m.group(2)
        if "_opaque" in body:
            continue
        fields = []
        for line in body.splitlines():
            fm = re.match(r"\s*pub (\w+): (.+),$", line)
            assert fm, line
            fields.append((fm.group(1), rust_type(fm.group(2))))
        structs[m.group(1)] = fields
    return structs


def c_params(params):
    params = params.strip()
    if params in ("", "void"):
        return []
    out = []
    for p in params.split(","):
        p = " ".join(p.split())
        pm = re.match(r"(.*?)(\w+)((?:\[\d+\])*)$", p)
        base, name, dims = pm.group(1), pm.group(2), pm.group(3)
        if dims:  # array parameters decay to pointers
            base = base + "*"
        out.append((name, c_type(base)))
    return out


def parse_c_functions(text):
    funcs = {}
    for m in re.finditer(r"^(tt_status|void|const char\*)\s+(tt_\w+)\s*\((.*?)\);", text, re.S | re.M):
        ret = None if m.group(1) == "void" else c_type(m.group(1))
        funcs[m.group(2)] = (ret, c_params(m.group(3)))
    return funcs


def parse_rust_functions(text):
    block = re.search(r'extern "C" \{(.*?)\n\}', text, re.S).group(1)
    funcs = {}
    for m in re.finditer(r"pub fn (\w+)\((.*?)\)(?: -> ([^;]+))?;", block, re.S):
        params = []
        for p in [x for x in m.group(2).split(",") if x.strip()]:
            name, ty = p.split(":", 1)
            params.append((name.strip(), rust_type(ty)))
        ret = rust_type(m.group(3)) if m.group(3) else None
        funcs[m.group(1)] = (ret, params)
    return funcs


def c_constants(text):
    consts = {}
    for m in re.finditer(r"\b(TT_\w+)\s*=\s*(\d+)", text):
        consts[m.group(1)] = int(m.group(2))
    for m in re.finditer(r"^#define (TT_\w+) (0x[0-9a-fA-F]+|\d+)u?$", text, re.M):
        consts[m.group(1)] = int(m.group(2), 0)
    return consts


def rust_constants(text):
    return {m.group(1): int(m.group(2).replace("_", ""), 0)
            for m in re.finditer(r"pub const (TT_\w+): \w+ = (0x[0-9a-fA-F_]+|\d+);", text)}


@pytest.fixture(scope="module")
def parsed():
    c_text = strip_c_comments(open(HEADER).read())
    r_text = open(CRATE).read()
    return c_text, r_text


def test_structs_have_the_same_fields_and_types(parsed):
    c_text, r_text = parsed
    c, r = parse_c_structs(c_text), parse_rust_structs(r_text)
    assert len(c) >= 30
    assert set(c) == set(r), (set(c) ^ set(r))
    for name in c:
        assert c[name] == r[name], name


def test_functions_have_the_same_signatures(parsed):
    c_text, r_text = parsed
    c, r = parse_c_functions(c_text), parse_rust_functions(r_text)
    assert set(c) == set(r), (set(c) ^ set(r))
    for name in c:
        c_ret, c_par = c[name]
        r_ret, r_par = r[name]
        assert c_ret == r_ret, name
        assert [t for _, t in c_par] == [t for _, t in r_par], name
        assert [n for n, _ in c_par] == [n for n, _ in r_par], name


def test_constants_have_the_same_values(parsed):
    c_text, r_text = parsed
    c, r = c_constants(c_text), rust_constants(r_text)
    assert set(c) == set(r), (set(c) ^ set(r))
    for name, value in c.items():
        assert r[name] == value, name


def rust_layout(structs, name):
    """Field offsets and the size of `name` under #[repr(C)]: each field at
    the next multiple of its alignment, size rounded up to the largest
    alignment."""
    def size_align(t):
        if isinstance(t, tuple) and t[0] == "ptr":
            return 8, 8
        if isinstance(t, tuple) and t[0] == "arr":
            s, a = size_align(t[1])
            return s * t[2], a
        if t in SIZES:
            return SIZES[t], SIZES[t]
        offsets, size, align = rust_layout(structs, t)
        return size, align

    offset, align, offsets = 0, 1, {}
    for field, t in structs[name]:
        s, a = size_align(t)
        offset = (offset + a - 1) // a * a
        offsets[field] = offset
        offset += s
        align = max(align, a)
    return offsets, (offset + align - 1) // align * align, align


def test_repr_c_layout_equals_gcc_offsets(parsed):
    c_text, r_text = parsed
    rust = parse_rust_structs(r_text)
    lines = ['#include <stddef.h>', '#include <stdio.h>', f'#include "{HEADER}"', "int main(void) {"]
    for name, fields in parse_c_structs(c_text).items():
        lines.append(f'printf("{name} . %zu\\n", sizeof({name}));')
        for field, _ in fields:
            lines.append(f'printf("{name} {field} %zu\\n", offsetof({name}, {field}));')
    lines += ["return 0;", "}"]
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "probe.c"), os.path.join(d, "probe")
        open(src, "w").write("\n".join(lines))
        subprocess.run(["gcc", "-std=c11", "-o", exe, src], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    gcc = {}
    for line in out.splitlines():
        s, f, v = line.split()
        gcc[(s, f)] = int(v)
    for name in rust:
        offsets, size, _ = rust_layout(rust, name)
        assert gcc[(name, ".")] == size, name
        for field, off in offsets.items():
            assert gcc[(name, field)] == off, (name, field)
    # The reference layouts keep their sizes.
    for name, size in (("tt_constants", 368), ("tt_vertex", 8), ("tt_mesh_record", 32),
                       ("tt_material_constants", 192), ("tt_material", 304), ("tt_instance", 208),
                       ("tt_emissive_triangle", 36)):
        assert gcc[(name, ".")] == size


def test_a_dropped_field_is_caught(parsed):
    """The comparison fails on the defect it exists for: a Rust tt_scene_desc
    that stops at emissive_triangle_count (112 B against 144 B)."""
    c_text, r_text = parsed
    truncated = re.sub(r"(pub emissive_triangle_count: u32,\n).*?(\n\})", r"\1    pub padding_: u32,\2",
                       r_text, count=1, flags=re.S)
    rust = parse_rust_structs(truncated)
    assert parse_c_structs(c_text)["tt_scene_desc"] != rust["tt_scene_desc"]
    _, size, _ = rust_layout(rust, "tt_scene_desc")
    assert size == 112
