"""Writes mojo/numax_hip.mojo, numax's side of the C ABI: one typed Mojo
wrapper per entry point declared in include/numax_hip.h, loaded through
`sys.ffi.DLHandle`.  Run after changing the header:

    python mojo/gen_numax_hip_mojo.py

tests/test_mojo_shim.py checks that the committed module is what this
script writes from the current header.  The image that builds this library
has no Mojo toolchain, so the module is not compiled here."""
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
HEADER = os.path.join(HERE, "..", "include", "numax_hip.h")
OUT = os.path.join(HERE, "numax_hip.mojo")

SCALARS = {"int": "Int32", "int64_t": "Int64", "uint64_t": "UInt64", "double": "Float64", "float": "Float32"}
POINTEES = {"char": "Int8", "int": "Int32", "double": "Float64", "float": "Float32", "int32_t": "Int32", "int64_t": "Int64", "uint8_t": "UInt8",
            "nmx_ir_cfg": "NmxIrCfg", "nmx_ir_report": "NmxIrReport", "nmx_batch_cfg": "NmxBatchCfg"}
OPAQUE = {"void", "nmx_ctx", "nmx_graph"}
STRUCTS = {"nmx_ir_cfg": "NmxIrCfg", "nmx_ir_report": "NmxIrReport", "nmx_batch_cfg": "NmxBatchCfg"}
RAW_INT = {"nmx_version"}  # int results that are values, not statuses
KEYWORDS = {"out", "ref", "mut", "owned", "inout", "read", "in", "fn", "var", "def"}  # C names Mojo reserves


def pname(name: str) -> str:
    return name + "_" if name in KEYWORDS else name


def mojo_type(ctype: str) -> str:
    t = ctype.replace("const ", "").strip()
    stars = t.count("*")
    base = t.replace("*", "").strip()
    if stars == 0:
        return SCALARS[base]
    if stars == 2:
        return "UnsafePointer[OpaquePointer]"
    if base in OPAQUE:
        return "OpaquePointer"
    return f"UnsafePointer[{POINTEES[base]}]"


def declarations(text: str):
    body = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    out = []
    for ret, name, args in re.findall(r"\b((?:const\s+)?(?:int|void|char)\s*\**)\s*(nmx_[a-z0-9_]+)\s*\(([^)]*)\)\s*;",
                                      body):
        params = []
        for a in args.split(","):
            a = " ".join(a.split())
            if not a or a == "void":
                continue
            m = re.match(r"(.*?)([A-Za-z_][A-Za-z0-9_]*)$", a)
            params.append((m.group(2), m.group(1).strip()))
        out.append((" ".join(ret.split()), name, params))
    return out


def structs(text: str):
    found = []
    for body, name in re.findall(r"typedef struct [a-z_]+ \{(.*?)\} ([a-z_]+);", text, flags=re.S):
        fields = []
        for line in re.sub(r"/\*.*?\*/", "", body, flags=re.S).split(";"):
            line = " ".join(line.split())
            if not line:
                continue
            m = re.match(r"([a-z0-9_]+) ([a-z_0-9]+)(?:\[(\d+)\])?$", line)
            fields.append((m.group(2), SCALARS[m.group(1)], m.group(3)))
        found.append((STRUCTS[name], name, fields))
    return found


def render(text: str) -> str:
    lines = [
        "# numax_hip.mojo -- numax's binding to libnumax_hip.so (include/numax_hip.h).",
        "#",
        "# Written by mojo/gen_numax_hip_mojo.py from the header; do not edit by hand.",
        "# One wrapper per C entry point, in header order: each takes the loaded",
        "# library first, then the C arguments with the same names and meaning, and",
        "# raises numax-style errors on a nonzero status (negative: argument -status",
        "# is invalid; positive: nmx_status_string's text).  Device pointers are the",
        "# unsafe pointers of numax's DeviceBuffers; the layouts are row-major with",
        "# explicit leading dimensions, as the header states per function.",
        "#",
        "# Not compiled in the image that builds the library (it has no Mojo",
        "# toolchain); tests/test_mojo_shim.py checks that every header entry point",
        "# has its wrapper here with the header's parameter list.",
        "",
        "from sys.ffi import DLHandle",
        "from memory import UnsafePointer, OpaquePointer",
        "",
        "",
        "struct NmxLib:",
        "    \"\"\"The loaded library (one per process).\"\"\"",
        "",
        "    var lib: DLHandle",
        "",
        "    fn __init__(out self, path: String = \"libnumax_hip.so\") raises:",
        "        self.lib = DLHandle(path)",
        "",
        "",
        "fn _check(h: NmxLib, status: Int32, what: String) raises:",
        "    if status == 0:",
        "        return",
        "    if status < 0:",
        "        raise Error(what + \": argument \" + String(-Int(status)) + \" is invalid\")",
        "    var text = h.lib.get_function[fn (Int32) -> UnsafePointer[UInt8]](\"nmx_status_string\")(status)",
        "    raise Error(what + \": \" + String(unsafe_from_utf8_ptr=text))",
        "",
    ]
    for mname, cname, fields in structs(text):
        lines += ["", "@register_passable(\"trivial\")", f"struct {mname}:",
                  f"    \"\"\"`{cname}` (same field order and types as the header).\"\"\"", ""]
        for fname, ftype, count in fields:
            if count:
                lines.append(f"    var {fname}: InlineArray[{ftype}, {count}]")
            else:
                lines.append(f"    var {fname}: {ftype}")
        lines.append("")
    for ret, name, params in declarations(text):
        sig = ", ".join(mojo_type(t) for _, t in params)
        args = ", ".join(f"{pname(p)}: {mojo_type(t)}" for p, t in params)
        call = ", ".join(pname(p) for p, _ in params)
        head = f"fn {name}(h: NmxLib" + (f", {args}" if args else "") + ")"
        lines.append("")
        if ret == "void":
            lines += [head + " raises:",
                      f"    var f = h.lib.get_function[fn ({sig}) -> None](\"{name}\")",
                      f"    f({call})"]
        elif ret == "const char*":
            lines += [head + " raises -> String:",
                      f"    var f = h.lib.get_function[fn ({sig}) -> UnsafePointer[UInt8]](\"{name}\")",
                      f"    return String(unsafe_from_utf8_ptr=f({call}))"]
        elif name in RAW_INT:
            lines += [head + " raises -> Int32:",
                      f"    var f = h.lib.get_function[fn ({sig}) -> Int32](\"{name}\")",
                      f"    return f({call})"]
        else:
            lines += [head + " raises:",
                      f"    var f = h.lib.get_function[fn ({sig}) -> Int32](\"{name}\")",
                      f"    _check(h, f({call}), \"{name}\")"]
    return "\n".join(lines) + "\n"


if __name__ == "__main__":
    src = render(open(HEADER).read())
    if len(sys.argv) > 1 and sys.argv[1] == "--check":
        sys.exit(0 if open(OUT).read() == src else 1)
    open(OUT, "w").write(src)
    print(f"wrote {OUT}")
