"""Writes tests/golden/reference_abi.json: the reference library's C ABI as data.

Three tables, read as text from the reference tree (nothing of it is imported or run):

* ``c_declarations`` -- every ``extern "C"`` entry point of src/petl_c_interface.h: return type, argument
  types (normalised: ``void bool int float char* void* int* float* float**``, names and ``const``
  dropped) and line;
* ``ctypes_bindings`` -- every ``restype`` / ``argtypes`` the reference's Python client
  (src/petl_server.py) declares for those entry points, in the same vocabulary (``POINTER(c_float)`` is
  ``float*``, ``POINTER(POINTER(c_float))`` is ``float**``) plus ``float[]`` / ``int[]`` for
  ``ndpointer(c_float)`` / ``ndpointer(c_int)`` (a C-contiguous numpy array passed as that pointer), each
  distinct argument list once, and the lines they sit on;
* ``calls`` -- per entry point, the lines of src/petl_server.py that call it.

The tests bind this build's libpetl.so exactly as that client does and check the declarations of
include/petl_c_interface.h against the first table.

    python tests/golden/make_reference_abi.py [--reference /root/reference]
"""
import argparse
import json
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "reference_abi.json")

C_TYPES = {"void": "void", "bool": "bool", "int": "int", "float": "float", "char*": "char*", "void*": "void*",
           "int*": "int*", "float*": "float*", "float**": "float**"}


def c_type(decl):
    """'const float* x' / 'float**' / 'int which' -> normalised type."""
    decl = decl.replace("const ", " ").strip()
    stars = decl.count("*")
    base = decl.replace("*", " ").split()[0]
    t = base + "*" * stars
    if t not in C_TYPES:
        raise ValueError(f"unexpected C type {decl!r}")
    return t


def parse_header(text):
    out = {}
    pat = re.compile(r'extern\s+"C"\s+\w+\s+([\w\*\s]+?)\s*\b(\w+)\s*\(([^)]*)\)\s*;')
    for lineno, line in enumerate(text.splitlines(), 1):
        if line.strip().startswith("//"):
            continue
        m = pat.search(line)
        if not m:
            continue
        ret, name, args = m.group(1), m.group(2), m.group(3).strip()
        argl = [] if args in ("", "void") else [c_type(a) for a in args.split(",")]
        out[name] = {"ret": c_type(ret), "args": argl, "line": lineno}
    return out


def split_top(s):
    parts, depth, cur = [], 0, ""
    for ch in s:
        if ch == "(":
            depth += 1
        elif ch == ")":
            depth -= 1
        if ch == "," and depth == 0:
            parts.append(cur.strip())
            cur = ""
        else:
            cur += ch
    if cur.strip():
        parts.append(cur.strip())
    return parts


def ctypes_type(expr):
    e = expr.replace(" ", "")
    table = {
        "ctypes.c_bool": "bool", "ctypes.c_int": "int", "ctypes.c_float": "float", "ctypes.c_char_p": "char*",
        "ctypes.c_void_p": "void*",
        'ndpointer(ctypes.c_float,flags="C_CONTIGUOUS")': "float[]",
        'ndpointer(ctypes.c_int,flags="C_CONTIGUOUS")': "int[]",
        "ctypes.POINTER(ctypes.c_float)": "float*",
        "ctypes.POINTER(ctypes.POINTER(ctypes.c_float))": "float**",
        "ctypes.POINTER(c_float_p)": "float**",
    }
    if e not in table:
        raise ValueError(f"unexpected ctypes type {expr!r}")
    return table[e]


def parse_client(text, names):
    bind, calls = {}, {}
    res = re.compile(r"libprojectors\.(\w+)\.restype\s*=\s*(\S+)")
    arg = re.compile(r"libprojectors\.(\w+)\.argtypes\s*=\s*\[(.*)\]\s*$")
    call = re.compile(r"libprojectors\.(\w+)\(")
    for lineno, line in enumerate(text.splitlines(), 1):
        s = line.strip()
        if s.startswith("#"):
            continue
        m = res.search(s)
        if m and m.group(1) in names:
            b = bind.setdefault(m.group(1), {"restype": None, "argtypes": [], "lines": []})
            t = ctypes_type(m.group(2))
            if b["restype"] not in (None, t):
                raise ValueError(f"{m.group(1)}: two restypes")
            b["restype"] = t
            b["lines"].append(lineno)
            continue
        m = arg.search(s)
        if m and m.group(1) in names:
            b = bind.setdefault(m.group(1), {"restype": None, "argtypes": [], "lines": []})
            a = [ctypes_type(x) for x in split_top(m.group(2))]
            if a not in b["argtypes"]:
                b["argtypes"].append(a)
            b["lines"].append(lineno)
            continue
        for m in call.finditer(s):
            if m.group(1) in names:
                calls.setdefault(m.group(1), []).append(lineno)
    return bind, calls


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--reference", default="/root/reference")
    a = ap.parse_args()
    with open(os.path.join(a.reference, "src", "petl_c_interface.h")) as fh:
        decls = parse_header(fh.read())
    with open(os.path.join(a.reference, "src", "petl_server.py")) as fh:
        bind, calls = parse_client(fh.read(), set(decls))
    doc = {"header": "src/petl_c_interface.h", "client": "src/petl_server.py",
           "c_declarations": decls, "ctypes_bindings": bind, "calls": calls}
    with open(OUT, "w") as fh:
        json.dump(doc, fh, indent=1, sort_keys=True)
    print(f"{len(decls)} declarations, {len(bind)} bound, {len(calls)} called -> {OUT}")


if __name__ == "__main__":
    main()
