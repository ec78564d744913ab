"""The Rust FFI crate (crates/holomap-gpu-sys) against the C header it
binds (include/holomap_gpu.h): every function declared with the same
parameter count and a matching Rust type per parameter, every constant with
the same value, every struct with the same fields in the same order and
types, and every HIP source of the library in build.rs. (No Rust toolchain
is installed here, so the crate is checked as text.)"""
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CRATE = os.path.join(ROOT, "crates", "holomap-gpu-sys")

C2RUST = {
    "hm_status": "hm_status", "void": None, "int": "c_int", "int32_t": "i32", "uint32_t": "u32",
    "uint64_t": "u64", "float": "f32", "double": "f64", "const char*": "*const c_char",
}


def header():
    t = open(os.path.join(ROOT, "include", "holomap_gpu.h")).read()
    return re.sub(r"/\*.*?\*/", "", t, flags=re.S)


def rust():
    return open(os.path.join(CRATE, "src", "lib.rs")).read()


def c_type_to_rust(t):
    t = " ".join(t.replace("*", " * ").split())
    ptrs = t.count("*")
    base = t.replace("*", "").replace("const", "").strip()
    const_first = t.startswith("const")
    r = C2RUST.get(base, base)
    if ptrs == 0:
        return r
    if base == "char":
        return "*const c_char"
    if ptrs == 1:
        return ("*const " if const_first else "*mut ") + r
    # T** (handles and output pointer arrays)
    inner = ("*const " if const_first else "*mut ") + r
    return "*mut " + inner


def test_functions_match():
    h = header()
    protos = re.findall(r"\n(?:hm_status|void|const char\*|int|uint64_t|uint32_t)\s+(hm_\w+)"
                        r"\s*\(([^;]*)\);", h)
    r = rust()
    assert len(protos) >= 45
    for name, args in protos:
        m = re.search(r"pub fn " + name + r"\(([^;]*?)\)(?:\s*->\s*([\w\s\*]+))?;", r, flags=re.S)
        assert m, f"{name} missing from lib.rs"
        cargs = [a.strip() for a in args.split(",") if a.strip() and a.strip() != "void"]
        rargs = [a.strip() for a in m.group(1).split(",") if a.strip()]
        assert len(cargs) == len(rargs), name
        for ca, ra in zip(cargs, rargs):
            ctype = re.sub(r"\s*\w+\s*$", "", ca).strip() if " " in ca else ca
            rtype = ra.split(":", 1)[1].strip()
            assert " ".join(rtype.split()) == c_type_to_rust(ctype), (name, ca, ra)


def test_constants_match():
    h = header()
    r = rust()
    for name, val in re.findall(r"#define (HM_\w+) (.+)", h):
        val = val.strip()
        if name.startswith("HOLOMAP"):
            continue
        m = re.search(r"pub const " + name + r": \w+ = ([^;]+);", r)
        assert m, name
        cv = eval(val.replace("ull", "").replace("u", ""))
        assert eval(m.group(1)) == cv, name


def struct_fields(text, name, rust_side):
    if rust_side:
        m = re.search(r"pub struct " + name + r" \{(.*?)\n\}", text, flags=re.S)
        return [(f.strip(), t.strip()) for f, t in re.findall(r"pub (\w+):\s*([^,]+),", m.group(1))]
    m = re.search(r"typedef struct \{([^{}]*)\}\s*" + name + ";", text)
    out = []
    for line in m.group(1).split(";"):
        line = " ".join(line.split())
        if not line:
            continue
        t, f = line.rsplit(" ", 1)
        arr = re.fullmatch(r"(\w+)\[(\d+)\]", f)
        if arr:  # a fixed-size array field
            out.append((arr.group(1), f"[{c_type_to_rust(t)}; {arr.group(2)}]"))
            continue
        out.append((f.lstrip("*"), c_type_to_rust(t + "*" * f.count("*"))))
    return out


def test_structs_match():
    h = header()
    r = rust()
    for name in ["hm_options", "hm_graph_out", "hm_server_config"]:
        assert struct_fields(h, name, False) == struct_fields(r, name, True), name


def test_build_script_compiles_every_library_source():
    import importlib.util
    spec = importlib.util.spec_from_file_location("b", os.path.join(ROOT, "holomap_amd", "build.py"))
    b = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(b)
    text = open(os.path.join(CRATE, "build.rs")).read()
    names = re.findall(r'"(hm_\w+)"', text.split("const SOURCES")[1].split("];")[0])
    assert sorted(n + ".hip" for n in names) == sorted(b.SOURCES)
