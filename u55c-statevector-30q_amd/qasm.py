"""OpenQASM 2.0 reader and writer for the Qiskit-free front-end.

The reference feeds its compiler Qiskit circuits and keeps its benchmark instances as OpenQASM 2
files (circuits/qasm/*.qasm, written by circuits/generators.py through Qiskit's qasm2 exporter).
This module reads such files into `circuit.Circuit` gate lists -- so the reference's instances run
through compiler -> QSV2 records -> HIP engine with no Qiskit -- and writes Circuits back out.

Accepted: the OPENQASM 2.0 header, `include "qelib1.inc"`, any number of qreg / creg declarations
(quantum registers are laid out in declaration order, register bits little-endian), the qelib1
gate names below, register broadcasting (`h q;`, `cx a, b;` over equal-size registers), user
`gate` definitions (expanded in place), `barrier` and `measure` (no effect on the statevector; the
reference compiler drops them, host/sv/compiler.py:20).  Parameters are expressions over numbers,
pi, + - * / ^, unary minus, parentheses and sin cos tan exp ln sqrt.  `reset`, `opaque` and `if`
have no statevector-only meaning here and raise ValueError.
"""
import math
import re

from .circuit import Circuit

# qelib1.inc name -> (Circuit gate name, qubits, parameters)
_GATES = {
    "U": ("u", 1, 3), "u": ("u", 1, 3), "u3": ("u", 1, 3), "u2": ("u2", 1, 2), "u1": ("u1", 1, 1),
    "p": ("p", 1, 1), "phase": ("p", 1, 1), "u0": ("id", 1, 1), "id": ("id", 1, 0),
    "x": ("x", 1, 0), "y": ("y", 1, 0), "z": ("z", 1, 0), "h": ("h", 1, 0), "s": ("s", 1, 0),
    "sdg": ("sdg", 1, 0), "t": ("t", 1, 0), "tdg": ("tdg", 1, 0), "sx": ("sx", 1, 0),
    "sxdg": ("sxdg", 1, 0), "rx": ("rx", 1, 1), "ry": ("ry", 1, 1), "rz": ("rz", 1, 1),
    "r": ("r", 1, 2),
    "CX": ("cx", 2, 0), "cx": ("cx", 2, 0), "cy": ("cy", 2, 0), "cz": ("cz", 2, 0), "ch": ("ch", 2, 0),
    "swap": ("swap", 2, 0), "iswap": ("iswap", 2, 0), "ecr": ("ecr", 2, 0), "dcx": ("dcx", 2, 0),
    "csx": ("csx", 2, 0), "cs": ("cs", 2, 0), "csdg": ("csdg", 2, 0),
    "crx": ("crx", 2, 1), "cry": ("cry", 2, 1), "crz": ("crz", 2, 1), "cp": ("cp", 2, 1),
    "cphase": ("cp", 2, 1), "cu1": ("cu1", 2, 1), "rxx": ("rxx", 2, 1), "ryy": ("ryy", 2, 1),
    "rzz": ("rzz", 2, 1), "rzx": ("rzx", 2, 1), "cu3": ("cu", 2, 3), "cu": ("cu", 2, 4),
    "xx_plus_yy": ("xx_plus_yy", 2, 2),
    "ccx": ("ccx", 3, 0), "ccz": ("ccz", 3, 0), "cswap": ("cswap", 3, 0),
    "c3x": ("mcx", 4, 0), "c4x": ("mcx", 5, 0),
}

_FUNCS = {"sin": math.sin, "cos": math.cos, "tan": math.tan, "exp": math.exp, "ln": math.log,
          "sqrt": math.sqrt}
_TOKEN = re.compile(r"\s*(?:(\d+\.\d*(?:[eE][-+]?\d+)?|\.\d+(?:[eE][-+]?\d+)?|\d+(?:[eE][-+]?\d+)?)|"
                    r"([A-Za-z_][A-Za-z0-9_]*)|(\*\*|[-+*/^()]))")


def eval_expr(text, env=None):
    """Value of an OpenQASM 2 parameter expression (recursive descent; names from `env` or pi)."""
    env = env or {}
    toks = []
    pos = 0
    text = text.strip()
    while pos < len(text):
        m = _TOKEN.match(text, pos)
        if not m or m.end() == pos:
            raise ValueError(f"bad expression {text!r} at {pos}")
        num, name, op = m.groups()
        toks.append(("n", float(num)) if num else ("i", name) if name else ("o", "^" if op == "**" else op))
        pos = m.end()
    i = 0

    def peek():
        return toks[i] if i < len(toks) else (None, None)

    def take(expect=None):
        nonlocal i
        t = peek()
        if t[0] is None or (expect is not None and t[1] != expect):
            raise ValueError(f"bad expression {text!r}: expected {expect or 'a term'}")
        i += 1
        return t

    def expr():
        v = term()
        while peek() in (("o", "+"), ("o", "-")):
            v = v + term() if take()[1] == "+" else v - term()
        return v

    def term():
        v = unary()
        while peek() in (("o", "*"), ("o", "/")):
            v = v * unary() if take()[1] == "*" else v / unary()
        return v

    def unary():
        if peek() == ("o", "-"):
            take()
            return -unary()
        if peek() == ("o", "+"):
            take()
            return unary()
        return power()

    def power():
        b = atom()
        if peek() == ("o", "^"):
            take()
            return b ** unary()          # right-associative
        return b

    def atom():
        kind, val = take()
        if kind == "n":
            return val
        if kind == "o" and val == "(":
            v = expr()
            take(")")
            return v
        if kind == "i":
            if val in _FUNCS:
                take("(")
                v = expr()
                take(")")
                return _FUNCS[val](v)
            if val == "pi":
                return math.pi
            if val in env:
                return env[val]
            raise ValueError(f"unknown name {val!r} in {text!r}")
        raise ValueError(f"bad expression {text!r}")

    v = expr()
    if i != len(toks):
        raise ValueError(f"trailing tokens in expression {text!r}")
    return float(v)


def _split_top(s, sep=","):
    """Split on `sep` outside parentheses."""
    out, depth, cur = [], 0, []
    for ch in s:
        if ch == "(":
            depth += 1
        elif ch == ")":
            depth -= 1
        if ch == sep and depth == 0:
            out.append("".join(cur))
            cur = []
        else:
            cur.append(ch)
    if cur or out:
        out.append("".join(cur))
    return [x.strip() for x in out]


_STMT = re.compile(r"^([A-Za-z_][A-Za-z0-9_]*)\s*(?:\((.*)\))?\s*(.*)$", re.S)


def loads(text, name=""):
    """Circuit from OpenQASM 2.0 source text."""
    src = re.sub(r"//[^\n]*", "", text)
    # statements end with ';' except gate bodies, which are { ... } blocks
    stmts = []
    pos = 0
    while pos < len(src):
        while pos < len(src) and src[pos].isspace():
            pos += 1
        if pos >= len(src):
            break
        semi = src.find(";", pos)
        brace = src.find("{", pos)
        if brace != -1 and (semi == -1 or brace < semi):
            close = src.find("}", brace)
            if close == -1:
                raise ValueError("unterminated gate body")
            stmts.append(src[pos:close + 1].strip())
            pos = close + 1
        else:
            if semi == -1:
                if src[pos:].strip():
                    raise ValueError(f"statement without ';': {src[pos:].strip()[:40]!r}")
                break
            stmts.append(src[pos:semi].strip())
            pos = semi + 1
    regs = {}            # name -> (offset, size)
    n = 0
    defs = {}            # user gate name -> (param names, qubit arg names, body statements)
    body = []            # (gate, [qubit indices], [param values])

    def expand(gname, qubits, params, depth=0):
        if depth > 64:
            raise ValueError("gate definitions nest too deeply")
        if gname in defs:
            pnames, qnames, stmts_ = defs[gname]
            if len(pnames) != len(params) or len(qnames) != len(qubits):
                raise ValueError(f"gate {gname}: wrong number of arguments")
            env = dict(zip(pnames, params))
            qmap = dict(zip(qnames, qubits))
            for st in stmts_:
                m = _STMT.match(st)
                if not m:
                    raise ValueError(f"bad statement in gate {gname}: {st!r}")
                g2, ptxt, qtxt = m.groups()
                if g2 == "barrier":
                    continue
                ps = [eval_expr(p, env) for p in _split_top(ptxt)] if ptxt and ptxt.strip() else []
                qs = [qmap[a.strip()] for a in qtxt.split(",")]
                expand(g2, qs, ps, depth + 1)
            return
        if gname not in _GATES:
            raise ValueError(f"unsupported gate {gname!r}")
        cname, nq, npar = _GATES[gname]
        if len(qubits) != nq or len(params) != npar:
            raise ValueError(f"gate {gname}: expects {nq} qubits and {npar} parameters")
        if gname == "cu3":
            params = list(params) + [0.0]
        if gname == "u0":
            params = []
        body.append((cname, list(qubits), list(params)))

    def qarg(a):
        a = a.strip()
        m = re.match(r"^([A-Za-z_][A-Za-z0-9_]*)\s*(?:\[\s*(\d+)\s*\])?$", a)
        if not m or m.group(1) not in regs:
            raise ValueError(f"unknown qubit argument {a!r}")
        off, size = regs[m.group(1)]
        if m.group(2) is None:
            return [off + i for i in range(size)]
        idx = int(m.group(2))
        if idx >= size:
            raise ValueError(f"index {idx} out of range for register {m.group(1)}[{size}]")
        return [off + idx]

    for st in stmts:
        if st.startswith("OPENQASM"):
            if not re.match(r"^OPENQASM\s+2(\.0)?$", st):
                raise ValueError(f"only OpenQASM 2.0 is read: {st!r}")
            continue
        if st.startswith("include"):
            continue
        m = re.match(r"^(qreg|creg)\s+([A-Za-z_][A-Za-z0-9_]*)\s*\[\s*(\d+)\s*\]$", st)
        if m:
            if m.group(1) == "qreg":
                regs[m.group(2)] = (n, int(m.group(3)))
                n += int(m.group(3))
            continue
        if st.startswith("gate"):
            m = re.match(r"^gate\s+([A-Za-z_][A-Za-z0-9_]*)\s*(?:\(([^)]*)\))?\s*([^{]*)\{(.*)\}$", st, re.S)
            if not m:
                raise ValueError(f"bad gate definition {st[:60]!r}")
            pnames = [p.strip() for p in m.group(2).split(",")] if m.group(2) and m.group(2).strip() else []
            qnames = [q.strip() for q in m.group(3).split(",")]
            inner = [s.strip() for s in m.group(4).split(";") if s.strip()]
            defs[m.group(1)] = (pnames, qnames, inner)
            continue
        head = st.split(None, 1)[0].split("(")[0]
        if head in ("reset", "opaque", "if"):
            raise ValueError(f"'{head}' has no statevector-only meaning here")
        if head == "measure" or head == "barrier":
            continue
        m = _STMT.match(st)
        if not m:
            raise ValueError(f"bad statement {st!r}")
        gname, ptxt, qtxt = m.groups()
        params = [eval_expr(p) for p in _split_top(ptxt)] if ptxt and ptxt.strip() else []
        args = [qarg(a) for a in _split_top(qtxt)]
        width = max(len(a) for a in args)
        if any(len(a) not in (1, width) for a in args):
            raise ValueError(f"register sizes differ in {st!r}")
        for j in range(width):
            expand(gname, [a[0] if len(a) == 1 else a[j] for a in args], params)
    qc = Circuit(n, name)
    for cname, qs, ps in body:
        qc.append(cname, qs, ps)
    return qc


def load(path):
    """Circuit from an OpenQASM 2.0 file (named after the file)."""
    import os
    with open(path) as f:
        return loads(f.read(), os.path.splitext(os.path.basename(path))[0])


# Circuit gate name -> qelib1 spelling (and how many of its parameters it takes)
_OUT = {"u": "u3", "mcx": None, "mcp": None}


def dumps(qc):
    """OpenQASM 2.0 text of a Circuit (parameters written with repr: exact round trip)."""
    lines = ["OPENQASM 2.0;", 'include "qelib1.inc";', f"qreg q[{qc.num_qubits}];"]
    inv = {}
    for k, (cname, nq, npar) in _GATES.items():
        inv.setdefault((cname, nq), k)
    for name, qs, ps in qc.tuples():
        if name == "mcx":
            gname = {4: "c3x", 5: "c4x"}.get(len(qs))
            if gname is None and len(qs) == 3:
                gname = "ccx"
            if gname is None:
                raise ValueError(f"mcx on {len(qs)} qubits has no OpenQASM 2 qelib1 spelling")
        elif name == "cu" and len(ps) == 4 and ps[3] == 0.0:
            gname, ps = "cu3", ps[:3]
        else:
            gname = _OUT.get(name, name)
            if gname is None or (gname, len(qs)) not in {(k, v[1]) for k, v in _GATES.items()}:
                gname = inv.get((name, len(qs)))
            if gname is None:
                raise ValueError(f"gate {name} on {len(qs)} qubits has no OpenQASM 2 spelling here")
        ptxt = "(" + ",".join(repr(float(p)) for p in ps) + ")" if ps else ""
        lines.append(f"{gname}{ptxt} " + ",".join(f"q[{q}]" for q in qs) + ";")
    return "\n".join(lines) + "\n"
