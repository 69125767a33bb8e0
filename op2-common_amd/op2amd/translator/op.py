"""The OP2 program model the translator works on: loops, their arguments and
the application's constants, read from C++ sources.

The reference's translator-v2 builds the same model with libclang
(translator-v2/op2-translator/cpp/parser.py:90-240, op.py); this one reads the
call sites lexically, which is enough for OP2's fixed call forms."""
import ast
import os
import re
from dataclasses import dataclass, field
from typing import List, Optional

from . import source

ACCESS = {"OP_READ": 0, "OP_WRITE": 1, "OP_RW": 2, "OP_INC": 3, "OP_MIN": 4, "OP_MAX": 5}
TYPES = {"double": 8, "float": 4, "int": 4, "unsigned int": 4, "long long": 8, "bool": 1,
         "char": 1}


class OpError(Exception):
    pass


@dataclass
class ArgDat:
    dat: str
    idx: int
    map: Optional[str]  # None for OP_ID
    dim: int
    typ: str
    acc: str
    opt: Optional[str] = None

    @property
    def direct(self):
        return self.map is None


@dataclass
class ArgGbl:
    ptr: str
    dim: int
    typ: str
    acc: str
    opt: Optional[str] = None


@dataclass
class ArgIdx:
    idx: int
    map: Optional[str]


@dataclass
class ArgInfo:
    ptr: str
    dim: int
    typ: str
    ref: int  # index of the OP_MIN/OP_MAX gbl arg it reports for


@dataclass
class Loop:
    index: int  # 1-based within its program, as translator-v2 numbers loops
    kernel: str
    args: list
    stem: str

    @property
    def stub(self):
        return f"op_par_loop_{self.stem}_{self.index}_{self.kernel}"

    def indirect_writes(self):
        return [a for a in self.args if isinstance(a, ArgDat) and not a.direct
                and a.acc in ("OP_WRITE", "OP_RW")]

    def indirect_incs(self):
        return [a for a in self.args if isinstance(a, ArgDat) and not a.direct
                and a.acc == "OP_INC"]


@dataclass
class Const:
    ptr: str
    dim: int
    typ: str


@dataclass
class Program:
    path: str
    text: str
    stem: str
    loops: List[Loop] = field(default_factory=list)
    consts: List[Const] = field(default_factory=list)  # one per name
    const_calls: List[Const] = field(default_factory=list)  # one per call
    kernel_text: dict = field(default_factory=dict)
    # every function the device needs, user kernels and the functions of the
    # program they call, callees before callers
    device_text: dict = field(default_factory=dict)


def _type(s):
    t = source.string_literal(s)
    base = t.split(":")[0].strip()
    if base not in TYPES:
        raise OpError(f"unsupported type {t!r}")
    return base


def _int(s, consts, what):
    """An integer constant expression: literals, the program's integer
    constants and macros, + - * / % and parentheses (C semantics: division
    truncates toward zero), as in op_decl_dat(edges, map_dim + 1, ...)."""
    s = s.strip()
    if re.fullmatch(r"-?\d+", s):
        return int(s)
    if s in consts:
        return consts[s]
    try:
        v = _eval_int(ast.parse(s, mode="eval").body, consts)
    except (SyntaxError, ValueError, ZeroDivisionError):
        v = None
    if v is None:
        raise OpError(f"{what} must be an integer constant, got {s!r}")
    return v


def _eval_int(node, consts):
    if isinstance(node, ast.Constant) and type(node.value) is int:
        return node.value
    if isinstance(node, ast.Name):
        if node.id not in consts:
            raise ValueError(node.id)
        return consts[node.id]
    if isinstance(node, ast.UnaryOp) and isinstance(node.op, (ast.USub, ast.UAdd)):
        v = _eval_int(node.operand, consts)
        return -v if isinstance(node.op, ast.USub) else v
    if isinstance(node, ast.BinOp):
        a, b = _eval_int(node.left, consts), _eval_int(node.right, consts)
        if isinstance(node.op, ast.Add):
            return a + b
        if isinstance(node.op, ast.Sub):
            return a - b
        if isinstance(node.op, ast.Mult):
            return a * b
        if isinstance(node.op, (ast.Div, ast.FloorDiv)):
            q = abs(a) // abs(b)
            return q if (a >= 0) == (b >= 0) else -q
        if isinstance(node.op, ast.Mod):
            r = abs(a) % abs(b)
            return r if a >= 0 else -r
    raise ValueError("not an integer constant expression")


def _map(s):
    s = s.strip()
    return None if s == "OP_ID" else s


def parse_arg(expr, consts):
    m = re.fullmatch(r"\s*(op_arg_dat|op_opt_arg_dat|op_arg_gbl|op_opt_arg_gbl|op_arg_idx|"
                     r"op_arg_info)\s*\((.*)\)\s*", expr, re.S)
    if not m:
        raise OpError(f"unrecognised op_par_loop argument: {expr.strip()[:80]}")
    kind, a = m.group(1), source.split_args(m.group(2))
    if kind == "op_arg_dat":
        if len(a) != 6:
            raise OpError("op_arg_dat takes 6 arguments")
        return ArgDat(a[0], _int(a[1], consts, "idx"), _map(a[2]), _int(a[3], consts, "dim"),
                      _type(a[4]), a[5].strip())
    if kind == "op_opt_arg_dat":
        if len(a) != 7:
            raise OpError("op_opt_arg_dat takes 7 arguments")
        return ArgDat(a[1], _int(a[2], consts, "idx"), _map(a[3]), _int(a[4], consts, "dim"),
                      _type(a[5]), a[6].strip(), opt=a[0])
    if kind == "op_arg_gbl":
        if len(a) != 4:
            raise OpError("op_arg_gbl takes 4 arguments")
        return ArgGbl(a[0], _int(a[1], consts, "dim"), _type(a[2]), a[3].strip())
    if kind == "op_opt_arg_gbl":
        if len(a) != 5:
            raise OpError("op_opt_arg_gbl takes 5 arguments")
        return ArgGbl(a[1], _int(a[2], consts, "dim"), _type(a[3]), a[4].strip(), opt=a[0])
    if kind == "op_arg_idx":
        return ArgIdx(_int(a[0], consts, "idx"), _map(a[1]))
    return ArgInfo(a[0], _int(a[1], consts, "dim"), _type(a[2]), _int(a[3], consts, "ref"))


_NOT_CALLS = {"if", "for", "while", "switch", "return", "sizeof", "catch"}


def _called_names(fn_text):
    """Names called in a function body, in order of first call; member calls
    (a.f(), p->f()) and qualified ones (std::f()) are left out."""
    body = source.strip_comments(fn_text[fn_text.index("{"):])
    out = []
    for m in re.finditer(r"(?<![\w.:>])([A-Za-z_]\w*)\s*\(", body):
        if m.group(1) not in _NOT_CALLS and m.group(1) not in out:
            out.append(m.group(1))
    return out


def _kernel_source(name, text, path, seen=None):
    """Definition text of user kernel `name`: in the file or a quoted include."""
    span = source.find_function(text, name)
    if span:
        return text[span[0]:span[1]]
    seen = seen or set()
    for inc in re.findall(r'^\s*#\s*include\s+"([^"]+)"', source.strip_comments(text), re.M):
        p = os.path.join(os.path.dirname(path), inc)
        if p in seen or not os.path.exists(p):
            continue
        seen.add(p)
        with open(p) as f:
            found = _kernel_source(name, f.read(), p, seen)
        if found:
            return found
    return None


def parse_program(path):
    with open(path) as f:
        text = f.read()
    clean = source.strip_comments(text)
    stem = os.path.splitext(os.path.basename(path))[0]
    prog = Program(path, text, stem)
    consts = source.int_constants(clean)
    for start, end, args in source.find_calls(clean, "op_decl_const"):
        if len(args) != 3:
            raise OpError("op_decl_const takes (dim, type, pointer)")
        ptr = re.sub(r"[&\s]|\[.*\]", "", args[2])
        c = Const(ptr, _int(args[0], consts, "dim"), _type(args[1]))
        prog.const_calls.append(c)
        same = [k for k in prog.consts if k.ptr == ptr]
        if same and (same[0].dim, same[0].typ) != (c.dim, c.typ):
            raise OpError(f"constant {ptr} declared with two shapes")
        if not same:
            prog.consts.append(c)
    for start, end, args in source.find_calls(clean, "op_par_loop"):
        if len(args) < 3:
            raise OpError("op_par_loop needs a kernel, a name and a set")
        kernel = args[0].strip()
        loop = Loop(len(prog.loops) + 1, kernel, [parse_arg(a, consts) for a in args[3:]], stem)
        for a in loop.args:
            if isinstance(a, ArgInfo) and not (0 <= a.ref < len(loop.args) and
                                               isinstance(loop.args[a.ref], ArgGbl) and
                                               loop.args[a.ref].acc in ("OP_MIN", "OP_MAX")):
                raise OpError(f"{kernel}: op_arg_info must refer to an OP_MIN/OP_MAX gbl arg")
        prog.loops.append(loop)
        if kernel not in prog.kernel_text:
            k = _kernel_source(kernel, text, path)
            if k is None:
                raise OpError(f"definition of user kernel {kernel} not found")
            prog.kernel_text[kernel] = k

    def visit(name, fn_text, stack):
        for callee in _called_names(fn_text):
            if callee in prog.device_text or callee in stack or callee == name:
                continue
            t = prog.kernel_text.get(callee) or _kernel_source(callee, text, path)
            if t is not None:
                visit(callee, t, stack | {name})
        prog.device_text.setdefault(name, fn_text)
    for name, fn_text in prog.kernel_text.items():
        visit(name, fn_text, set())
    return prog


def translate_program(prog, header_name):
    """The application source with each op_par_loop(kernel, ...) rewritten to
    its stub, op_decl_const(dim, type, &x) to op_decl_const2("x", ...), and the
    OP2 headers replaced (translator-v2/op2-translator/cpp/translator/program.py
    does the same rewrites)."""
    text = prog.text
    clean = source.strip_comments(text)
    edits = []
    for loop, (start, end, args) in zip(prog.loops, source.find_calls(clean, "op_par_loop")):
        open_pos = clean.index("(", start)
        rest = clean[open_pos + 1:end - 1]
        comma = rest.index(",")
        edits.append((start, open_pos + 1 + comma + 1, f"{loop.stub}("))
    for const, (start, end, args) in zip(prog.const_calls, source.find_calls(clean, "op_decl_const")):
        open_pos = clean.index("(", start)
        edits.append((start, open_pos + 1, f'op_decl_const2("{const.ptr}", '))
    for m in re.finditer(r'^[ \t]*#[ \t]*include[ \t]+"(op_seq\.h|op_lib_cpp\.h|op_profile\.h)"[^\n]*',
                         clean, re.M):
        repl = (f'#include "op2_hip.h"\n#include "{header_name}"'
                if m.group(1) != "op_profile.h" else "// op_profile.h: declared by op2_hip.h")
        edits.append((m.start(), m.end(), repl))
    out = text
    for start, end, repl in sorted(edits, reverse=True):
        out = out[:start] + repl + out[end:]
    return out
