"""The C-ABI boundary without a GPU: every function that include/*.h
declares is exported by libfmftsaxs_amd.so, and every header compiles on its
own as C11 and as C++ (no compute calls)."""
import ctypes as C
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")
LIB = os.path.join(ROOT, "libfmftsaxs_amd", "libfmftsaxs_amd.so")


def _headers():
    return sorted(f for f in os.listdir(INC) if f.endswith(".h"))


def _declared_functions(text):
    """Names of the functions a header declares at file scope."""
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    text = re.sub(r"//[^\n]*", " ", text)
    text = "\n".join(l for l in text.split("\n") if not l.lstrip().startswith("#"))
    # the extern "C" wrapper: its opening line and its lone closing brace
    text = re.sub(r'extern\s+"C"\s*\{', " ", text)
    text = "\n".join(l for l in text.split("\n") if l.strip() != "}")
    # drop struct/union/enum bodies and typedefs (their members are not functions)
    prev = None
    while prev != text:
        prev = text
        text = re.sub(r"\{[^{}]*\}", " ", text)
    names = set()
    for stmt in text.split(";"):
        s = " ".join(stmt.split())
        if not s or s.startswith("typedef"):
            continue
        m = re.match(r"^(?:[A-Za-z_][\w\s\*]*?[\s\*])([A-Za-z_]\w*)\s*\((.*)\)$", s)
        if m and "(*" not in s:
            names.add(m.group(1))
    return names


def test_headers_declare_functions():
    total = sum(len(_declared_functions(open(os.path.join(INC, h)).read())) for h in _headers())
    assert total > 60


@pytest.mark.parametrize("header", _headers())
def test_header_symbols_exported(header):
    if not os.path.exists(LIB):
        pytest.skip("library not built")
    lib = C.CDLL(LIB)
    names = _declared_functions(open(os.path.join(INC, header)).read())
    missing = sorted(n for n in names if not hasattr(lib, n))
    assert not missing, f"{header}: declared but not exported: {missing}"


@pytest.mark.parametrize("header", _headers())
@pytest.mark.parametrize("lang", ["c", "c++"])
def test_header_compiles_alone(header, lang, tmp_path):
    src = tmp_path / ("t.c" if lang == "c" else "t.cpp")
    src.write_text(f'#include "{header}"\nint main(void) {{ return 0; }}\n')
    cc = ["gcc", "-std=c11"] if lang == "c" else ["g++", "-std=c++17"]
    r = subprocess.run(cc + ["-fsyntax-only", "-Wall", "-I", INC, str(src)], capture_output=True,
                       text=True)
    assert r.returncode == 0, r.stderr
