"""Static reading of the Julia binding (gaugefields.jl_amd/julia/GaugefieldsHIP.jl) and of the C
header: every ccall's symbol, return type and argument types, and the ordered ccall sequence of
each Julia method (calls to other functions of the file expanded in place)."""
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
JULIA = os.path.join(ROOT, "gaugefields.jl_amd", "julia", "GaugefieldsHIP.jl")
HEADER = os.path.join(ROOT, "include", "gfhip.h")

# calls that are not part of a method's own sequence: the lazy gf_create of a context's first
# use and the error-message reads that only run after a failing status
NOT_IN_SEQUENCE = {"_ctx", "_check", "_check_ildg"}

# C parameter type -> kind
_C_KINDS = [
    (r"^(const\s+)?gf_ctx\s*\*\*$", "ptr*"),
    (r"^(const\s+)?gf_ctx\s*\*$", "ptr"),
    (r"^(const\s+)?void\s*\*$", "ptr"),
    (r"^(const\s+)?int\s*\*$", "int*"),
    (r"^(const\s+)?double\s*\*$", "double*"),
    (r"^(const\s+)?float\s*\*$", "float*"),
    (r"^(const\s+)?char\s*\*$", "cstr"),
    (r"^long\s+long\s*\*$", "i64*"),
    (r"^size_t\s*\*$", "size*"),
    (r"^int$", "int"),
    (r"^double$", "double"),
    (r"^unsigned\s+long\s+long$", "u64"),
    (r"^void$", "void"),
    (r"^const\s+char\s*\*$", "cstr"),
]

# Julia ccall type -> kind
_JL_KINDS = {
    "Ptr{Cvoid}": "ptr",
    "Ref{Ptr{Cvoid}}": "ptr*",
    "Ptr{Cint}": "int*",
    "Ref{Cint}": "int*",
    "Ptr{Float64}": "double*",
    "Ref{Float64}": "double*",
    "Ptr{ComplexF64}": "double*",
    "Ptr{Cfloat}": "float*",
    "Ref{Cfloat}": "float*",
    "Cstring": "cstr",
    "Ref{Clonglong}": "i64*",
    "Cint": "int",
    "Cdouble": "double",
    "Culonglong": "u64",
}


def _c_kind(t):
    t = " ".join(t.replace("*", " * ").split()).replace(" *", "*").replace("* ", "*")
    t = re.sub(r"\s*\*", "*", t)
    for pat, kind in _C_KINDS:
        if re.match(pat, t.replace("*", " *").replace("  ", " ").strip()) or re.match(pat, t):
            return kind
    raise ValueError(f"unmapped C type {t!r}")


def header_prototypes():
    """{symbol: (return kind, [parameter kinds])} of every function include/gfhip.h declares"""
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    out = {}
    for m in re.finditer(r"\b(int|const\s+char\s*\*)\s*(gf_\w+)\s*\(([^)]*)\)\s*;", text):
        ret, name, params = m.group(1), m.group(2), m.group(3).strip()
        kinds = []
        if params and params != "void":
            for p in params.split(","):
                p = p.strip()
                p = re.sub(r"\b\w+$", "", p).strip() if not p.endswith("*") else p  # drop the name
                kinds.append(_c_kind(p))
        out[name] = ("cstr" if "char" in ret else "int", kinds)
    return out


def _split_top(s):
    """split at top-level commas (outside {} and ())"""
    parts, depth, cur = [], 0, ""
    for ch in s:
        if ch in "({":
            depth += 1
        elif ch in ")}":
            depth -= 1
        if ch == "," and depth == 0:
            parts.append(cur)
            cur = ""
        else:
            cur += ch
    if cur.strip():
        parts.append(cur)
    return [p.strip() for p in parts if p.strip()]


def _matching(text, i):
    """index of the parenthesis closing the one at text[i]"""
    depth = 0
    for k in range(i, len(text)):
        if text[k] == "(":
            depth += 1
        elif text[k] == ")":
            depth -= 1
            if depth == 0:
                return k
    raise ValueError("unbalanced parentheses")


def julia_ccalls(text=None):
    """[(symbol, return type, [argument types], [argument expressions])] of every ccall"""
    text = open(JULIA).read() if text is None else text
    out = []
    for m in re.finditer(r"ccall\(", text):
        body = text[m.end(): _matching(text, m.end() - 1)]
        parts = _split_top(body)
        sym = re.match(r"\(:(\w+),\s*libgfhip\)", parts[0]).group(1)
        ret = parts[1]
        types = _split_top(parts[2].strip()[1:-1]) if parts[2].strip() != "()" else []
        out.append((sym, ret, types, parts[3:]))
    return out


def julia_kind(t):
    if t not in _JL_KINDS:
        raise ValueError(f"unmapped Julia ccall type {t!r}")
    return _JL_KINDS[t]


def _definitions(text):
    """[(name, signature text, body text)] of the file's function definitions, in order"""
    defs = []
    lines = text.split("\n")
    i = 0
    while i < len(lines):
        ln = lines[i]
        m = re.match(r"^function\s+((?:Base\.)?[\w!]+|Base\.:\*)\(", ln)
        if m:
            j = i + 1
            while lines[j] != "end":
                j += 1
            defs.append((m.group(1), ln, "\n".join(lines[i + 1: j])))
            i = j + 1
            continue
        m = re.match(r"^((?:Base\.)?[\w!]+)(\(.*)$", ln)
        if m and not ln.startswith(("module", "end", "const", "import", "using", "export", "struct")):
            # one-line definition "name(args...) = expr", possibly continued on indented lines
            j = i + 1
            while j < len(lines) and lines[j].startswith("    "):
                j += 1
            whole = "\n".join(lines[i:j])
            close = _matching(whole, len(m.group(1)))
            rest = whole[close + 1:]
            eq = re.match(r"^(\s*where\s*\{[^}]*\})?\s*=\s*", rest)
            if eq:
                defs.append((m.group(1), whole[: close + 1], rest[eq.end():]))
                i = j
                continue
        i += 1
    return defs


def _strip(code):
    """code with string literals blanked and comments removed (positions of the rest kept)"""
    out, i, n = [], 0, len(code)
    while i < n:
        ch = code[i]
        if ch == '"':
            j = i + 1
            while j < n and code[j] != '"':
                j += 2 if code[j] == "\\" else 1
            out.append('"' + " " * (j - i - 1) + '"')
            i = j + 1
        elif ch == "#":
            j = code.find("\n", i)
            j = n if j < 0 else j
            out.append(" " * (j - i))
            i = j
        else:
            out.append(ch)
            i += 1
    return "".join(out)


def method_sequences(text=None):
    """{signature text: [gf_ symbols in execution order]}: calls to functions of the file are
    expanded in place (a name with several methods expands as its first one), the arguments of
    a call run before it, NOT_IN_SEQUENCE is skipped"""
    text = open(JULIA).read() if text is None else text
    defs = [(n, sig, _strip(b)) for n, sig, b in _definitions(text)]
    first = {}
    for name, sig, body in defs:
        first.setdefault(name, body)
    names = sorted(first, key=len, reverse=True)
    call_re = re.compile(r"ccall\(\(:(gf_\w+)|(?<![\w.:])(" + "|".join(re.escape(n) for n in names) + r")\(")

    def seq(body, stack):
        # call sites with the span of their argument list; post-order over span containment
        sites = []
        for m in call_re.finditer(body):
            open_paren = m.start() + (len("ccall") if m.group(1) else len(m.group(2)))
            sites.append((m.start(), _matching(body, open_paren), m.group(1), m.group(2)))
        out, done = [], set()

        def emit(k):
            s0, e0, sym, name = sites[k]
            for j in range(k + 1, len(sites)):
                if sites[j][0] > e0:
                    break
                if j not in done:
                    emit(j)
            done.add(k)
            if sym:
                out.append(sym)
            elif name not in NOT_IN_SEQUENCE and name not in stack:
                out.extend(seq(first[name], stack | {name}))

        for k in range(len(sites)):
            if k not in done:
                emit(k)
        return out

    return {sig: seq(body, {name}) for name, sig, body in defs}


def sequence_of(selector, text=None):
    """the sequence of the one definition whose signature text contains `selector`"""
    seqs = method_sequences(text)
    hits = [s for s in seqs if selector in s]
    if len(hits) != 1:
        raise KeyError(f"{selector!r} selects {len(hits)} definitions: {hits}")
    return seqs[hits[0]]
