"""ctypes prototypes for every entry point of ``include/frovedis_hip.h``.

Without ``argtypes`` ctypes passes a Python integer as a C ``int``: a
``size_t`` above 2^31 or a 64-bit pointer then reaches the library with its
upper half undefined, which shows up as a host crash far from the call.  The
prototypes are read from the header itself, so the Python side can never
disagree with the C declarations; an argument that does not fit its C type
raises ``ctypes.ArgumentError`` at the call instead.
"""
import ctypes
import os
import re

HEADER = os.path.join(os.path.dirname(os.path.abspath(__file__)), os.pardir, "include",
                      "frovedis_hip.h")

_SCALARS = {
    "int": ctypes.c_int,
    "unsigned": ctypes.c_uint,
    "unsigned int": ctypes.c_uint,
    "size_t": ctypes.c_size_t,
    "long long": ctypes.c_longlong,
    "unsigned long long": ctypes.c_ulonglong,
    "int64_t": ctypes.c_int64,
    "uint64_t": ctypes.c_uint64,
    "float": ctypes.c_float,
    "double": ctypes.c_double,
}


def _ctype(decl, is_return):
    """The ctypes type of a C type (a parameter's name already removed)."""
    t = " ".join(decl.replace("*", " * ").split())
    if "*" in t:
        if is_return and t.replace("const ", "") == "char *":
            return ctypes.c_char_p
        return ctypes.c_void_p
    t = t.replace("const ", "")
    if t == "void" and is_return:
        return None
    if t not in _SCALARS:
        raise ValueError(f"no ctypes mapping for C type {decl!r}")
    return _SCALARS[t]


def _strip_name(param):
    """`const size_t* n` -> `const size_t*`; `size_t n` -> `size_t`."""
    p = param.strip()
    m = re.match(r"^(.*?[\s\*])([A-Za-z_]\w*)$", p)
    if not m:
        return p
    head = m.group(1).strip()
    # a lone type name (no parameter name) such as `size_t`
    if not head or head in ("const", "unsigned", "long", "signed"):
        return p
    return head


def prototypes(header=HEADER):
    """{name: (restype, [argtypes])} for every fhip_* declaration."""
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    text = "\n".join(line for line in text.splitlines() if not line.lstrip().startswith("#"))
    text = text.replace('extern "C" {', "").replace("}", " } ")
    out = {}
    for ret, name, params in re.findall(
            r"([A-Za-z_][\w\s\*]*?)\b(fhip_\w+)\s*\(([^;{}()]*)\)\s*;", text, flags=re.S):
        ret = ret.split(";")[-1].split("}")[-1]
        ret = re.sub(r"\btypedef\b.*", "", ret).strip()
        ps = [p for p in params.split(",") if p.strip()]
        if len(ps) == 1 and ps[0].strip() == "void":
            ps = []
        out[name] = (_ctype(ret, True), [_ctype(_strip_name(p), False) for p in ps])
    return out


def apply(lib, header=HEADER):
    """Sets restype and argtypes of every declared entry point on `lib`."""
    for name, (res, args) in prototypes(header).items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib
