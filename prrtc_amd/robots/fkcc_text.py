"""Read the reference's generated forward-kinematics + collision code as data.

The reference ships its robot models only as generated straight-line device
code: one ``fkcc<Robot>`` specialisation per robot in
src/planning/utils.cuh (Sphere :330-334, Panda :336-5270, Fetch :5272-19882,
Baxter :19884-33131).  There is no URDF and no sphere file.  This module
parses that text into a small program representation:

* ``assign``  -- ``float|auto NAME = EXPR;`` (an SSA value),
* ``check``   -- ``if (<call>) { ... }`` blocks whose condition is
  ``sphere_environment_in_collision(environment, x, y, z, r)`` or
  ``sphere_sphere_self_collision(ax, ay, az, ar, bx, by, bz, br)`` and whose
  body is either ``return false;`` or a list of nested leaf checks.

and evaluates the SSA values with the C++ type rules the generated code has
(``float`` temporaries round to binary32; ``auto`` with a ``double`` literal
promotes to binary64, utils.cuh:5363), vectorised over many poses with numpy.

It is used offline, once, to fit the kinematic tables in ``*.json`` and to
write golden vectors; nothing at run time reads the reference.
"""
from __future__ import annotations

import re
from dataclasses import dataclass, field

import numpy as np

ROBOT_MARKERS = {
    "panda": "fkcc<ppln::robots::Panda>",
    "fetch": "fkcc<ppln::robots::Fetch>",
    "baxter": "fkcc<ppln::robots::Baxter>",
}

_TOKEN_RE = re.compile(
    r"\s*(?:(?P<num>\d+\.\d*(?:[eE][-+]?\d+)?f?|\d+(?:[eE][-+]?\d+)?f?)|"
    r"(?P<name>[A-Za-z_][A-Za-z_0-9:]*)|(?P<op>[-+*/(),\[\];={}!<>]))"
)


@dataclass
class Expr:
    kind: str                 # 'num' | 'var' | 'input' | 'bin' | 'neg' | 'call'
    value: object = None      # number / name / input index / op / func name
    args: tuple = ()
    is_double: bool = False   # literal type (double unless 'f' suffix)


@dataclass
class Call:
    func: str                 # 'env' | 'self'
    args: tuple               # Expr per argument (environment argument dropped)
    label: str | None = None  # the /*link*/ comment in front of the call


@dataclass
class Check:
    call: Call
    leaves: list = field(default_factory=list)  # list[Call]; empty => direct


@dataclass
class Program:
    robot: str
    assigns: list             # (name, decl, Expr) in source order
    checks: list              # Check in source order
    order: list               # ('assign', i) / ('check', i) interleaving


def _read_body(text: str, marker: str) -> str:
    start = text.index(marker)
    brace = text.index("{", text.index(")", start))
    depth = 0
    for pos in range(brace, len(text)):
        ch = text[pos]
        if ch == "{":
            depth += 1
        elif ch == "}":
            depth -= 1
            if depth == 0:
                return text[brace + 1:pos]
    raise ValueError("unbalanced braces for " + marker)


class _Parser:
    def __init__(self, body: str):
        # Keep /*label*/ comments that sit directly inside an if condition by
        # turning them into a LABEL token; drop every other comment.
        body = re.sub(r"/\*\s*([A-Za-z0-9_. ]+?)\s*\*/", lambda m: " @LABEL[" + m.group(1).replace(" ", "~") + "] ", body)
        body = re.sub(r"//[^\n]*", "", body)
        self.toks = []
        pos = 0
        label_re = re.compile(r"\s*@LABEL\[([^\]]*)\]")
        while pos < len(body):
            m = label_re.match(body, pos)
            if m:
                self.toks.append(("label", m.group(1).replace("~", " ")))
                pos = m.end()
                continue
            m = _TOKEN_RE.match(body, pos)
            if not m or m.end() == pos:
                if body[pos:].strip() == "":
                    break
                raise ValueError("cannot tokenise near: " + body[pos:pos + 60])
            pos = m.end()
            if m.group("num"):
                self.toks.append(("num", m.group("num")))
            elif m.group("name"):
                self.toks.append(("name", m.group("name")))
            else:
                self.toks.append(("op", m.group("op")))
        self.i = 0

    def peek(self, k=0):
        return self.toks[self.i + k] if self.i + k < len(self.toks) else (None, None)

    def take(self, kind=None, value=None):
        tok = self.toks[self.i]
        if kind and tok[0] != kind or value and tok[1] != value:
            raise ValueError(f"expected {kind} {value}, got {tok} at token {self.i}")
        self.i += 1
        return tok

    # expression grammar: sum := prod (('+'|'-') prod)* ; prod := unary (('*'|'/') unary)*
    def expr(self):
        lhs = self.prod()
        while self.peek() in (("op", "+"), ("op", "-")):
            op = self.take()[1]
            lhs = Expr("bin", op, (lhs, self.prod()))
        return lhs

    def prod(self):
        lhs = self.unary()
        while self.peek() in (("op", "*"), ("op", "/")):
            op = self.take()[1]
            lhs = Expr("bin", op, (lhs, self.unary()))
        return lhs

    def unary(self):
        if self.peek() == ("op", "-"):
            self.take()
            return Expr("neg", None, (self.unary(),))
        if self.peek() == ("op", "+"):
            self.take()
            return self.unary()
        return self.atom()

    def atom(self):
        kind, val = self.peek()
        if kind == "num":
            self.take()
            is_f = val.endswith("f")
            return Expr("num", float(val.rstrip("f")), is_double=not is_f)
        if kind == "op" and val == "(":
            self.take()
            e = self.expr()
            self.take("op", ")")
            return e
        if kind == "name":
            self.take()
            if val == "q":
                self.take("op", "[")
                idx = int(self.take("num")[1])
                self.take("op", "]")
                return Expr("input", idx)
            if self.peek() == ("op", "("):
                self.take()
                args = [self.expr()]
                while self.peek() == ("op", ","):
                    self.take()
                    args.append(self.expr())
                self.take("op", ")")
                return Expr("call", val, tuple(args))
            return Expr("var", val)
        raise ValueError(f"unexpected token {kind} {val}")

    def call(self):
        label = None
        if self.peek()[0] == "label":
            label = self.take()[1]
        fname = self.take("name")[1]
        self.take("op", "(")
        args = []
        if fname == "sphere_environment_in_collision":
            self.take("name", "environment")
            self.take("op", ",")
            func = "env"
        elif fname == "sphere_sphere_self_collision":
            func = "self"
        else:
            raise ValueError("unknown call " + fname)
        args.append(self.expr())
        while self.peek() == ("op", ","):
            self.take()
            args.append(self.expr())
        self.take("op", ")")
        return Call(func, tuple(args), label)

    def return_false(self):
        self.take("name", "return")
        self.take("name", "false")
        self.take("op", ";")

    def program(self, robot: str) -> Program:
        assigns, checks, order = [], [], []
        while self.i < len(self.toks):
            kind, val = self.peek()
            if kind == "name" and val in ("float", "auto", "double"):
                self.take()
                name = self.take("name")[1]
                self.take("op", "=")
                e = self.expr()
                self.take("op", ";")
                order.append(("assign", len(assigns)))
                assigns.append((name, val, e))
            elif kind == "name" and val == "if":
                self.take()
                self.take("op", "(")
                c = self.call()
                self.take("op", ")")
                self.take("op", "{")
                leaves = []
                if self.peek() == ("name", "return"):
                    self.return_false()
                else:
                    while self.peek() == ("name", "if"):
                        self.take()
                        self.take("op", "(")
                        leaves.append(self.call())
                        self.take("op", ")")
                        self.take("op", "{")
                        self.return_false()
                        self.take("op", "}")
                self.take("op", "}")
                order.append(("check", len(checks)))
                checks.append(Check(c, leaves))
            elif kind == "name" and val == "return":
                self.take()
                self.take("name", "true")
                self.take("op", ";")
            else:
                raise ValueError(f"unexpected statement start {kind} {val} at {self.i}")
        return Program(robot, assigns, checks, order)


def parse(utils_cuh_path: str, robot: str) -> Program:
    with open(utils_cuh_path) as fh:
        text = fh.read()
    body = _read_body(text, ROBOT_MARKERS[robot])
    return _Parser(body).program(robot)


# ----------------------------------------------------------------------------
# evaluation

class Evaluator:
    """Evaluate the SSA values of a parsed program on a batch of poses.

    ``mode='ref'`` follows the generated code's own types (binary32 for
    ``float`` declarations and float-only ``auto`` expressions, binary64 where
    a double literal enters).  ``mode='f64'`` evaluates every value in binary64
    (used to fit the kinematic tables).
    """

    def __init__(self, prog: Program, q: np.ndarray, mode: str = "ref"):
        self.prog = prog
        self.mode = mode
        self.q = np.asarray(q, dtype=np.float32)
        self.vals = {}
        self.types = {}
        for name, decl, e in prog.assigns:
            v, is_d = self.ev(e)
            if mode == "ref" and decl == "float":
                v, is_d = v.astype(np.float32) if isinstance(v, np.ndarray) else np.float32(v), False
            self.vals[name] = v
            self.types[name] = is_d

    def ev(self, e: Expr):
        f64 = self.mode == "f64"
        if e.kind == "num":
            return (np.float64(e.value) if (e.is_double or f64) else np.float32(e.value)), (e.is_double or f64)
        if e.kind == "input":
            col = self.q[:, e.value]
            return (col.astype(np.float64) if f64 else col), f64
        if e.kind == "var":
            return self.vals[e.value], self.types[e.value]
        if e.kind == "neg":
            v, d = self.ev(e.args[0])
            return -v, d
        if e.kind == "bin":
            a, da = self.ev(e.args[0])
            b, db = self.ev(e.args[1])
            d = da or db
            dt = np.float64 if d else np.float32
            a = np.asarray(a, dtype=dt)
            b = np.asarray(b, dtype=dt)
            op = e.value
            r = a + b if op == "+" else a - b if op == "-" else a * b if op == "*" else a / b
            return r.astype(dt), d
        if e.kind == "call":
            v, d = self.ev(e.args[0])
            fn = {"sin": np.sin, "cos": np.cos, "sqrt": np.sqrt}[e.value]
            dt = np.float64 if d else np.float32
            return fn(np.asarray(v, dtype=dt)).astype(dt), d
        raise ValueError(e.kind)

    def arg(self, e: Expr, n: int, dtype=np.float32) -> np.ndarray:
        v, _ = self.ev(e)
        return np.broadcast_to(np.asarray(v, dtype=np.float64).astype(dtype), (n,)).copy()


def sphere_key(args) -> tuple:
    """Identity of a sphere argument quadruple (x, y, z, r) as text."""
    def k(e: Expr):
        if e.kind == "num":
            return repr(float(e.value))
        if e.kind == "var":
            return e.value
        if e.kind == "neg" and e.args[0].kind == "num":
            return repr(-float(e.args[0].value))
        raise ValueError("unsupported sphere argument " + str(e))
    return tuple(k(a) for a in args)


def dependencies(prog: Program) -> dict:
    """Map every SSA name to the frozenset of joint indices it depends on."""
    deps = {}

    def walk(e: Expr):
        if e.kind == "input":
            return frozenset([e.value])
        if e.kind == "var":
            return deps[e.value]
        out = frozenset()
        for a in e.args:
            out |= walk(a)
        return out

    for name, _decl, e in prog.assigns:
        deps[name] = walk(e)
    prog._walk = walk
    return deps


def structure(prog: Program):
    """Spheres (first-use order) and the check tables of a parsed program.

    Returns ``(spheres, env_checks, self_checks, order)`` where each sphere is
    ``{"key", "args", "label"}`` and the checks refer to sphere indices.
    """
    spheres, index = [], {}

    def sid(args, label=None):
        key = sphere_key(args)
        if key not in index:
            index[key] = len(spheres)
            spheres.append({"key": key, "args": args, "label": label})
        elif label and not spheres[index[key]]["label"]:
            spheres[index[key]]["label"] = label
        return index[key]

    env_checks, self_checks, order = [], [], []
    for kind, ci in prog.order:
        if kind != "check":
            continue
        chk = prog.checks[ci]
        c = chk.call
        if c.func == "env":
            g = sid(c.args, c.label)
            leaves = [sid(l.args, c.label) for l in chk.leaves]
            env_checks.append({"label": c.label or "", "gate": g, "leaves": leaves, "direct": not chk.leaves})
            order.append(["env", len(env_checks) - 1])
        else:
            la, lb = (c.label.split(" vs. ") + [None, None])[:2] if c.label else (None, None)
            ga, gb = sid(c.args[:4], la), sid(c.args[4:], lb)
            pairs = [[sid(l.args[:4], la), sid(l.args[4:], lb)] for l in chk.leaves]
            self_checks.append({"label": c.label or "", "gate": [ga, gb], "pairs": pairs, "direct": not chk.leaves})
            order.append(["self", len(self_checks) - 1])
    return spheres, env_checks, self_checks, order


def sphere_centres(prog: Program, spheres, q: np.ndarray, mode: str = "ref") -> np.ndarray:
    """Centres (n, S, 3) of the given spheres at poses q.

    In ``ref`` mode the result is what the generated code passes to the
    collision routines: each coordinate narrowed to binary32 at the call
    (their parameters are ``float``, utils.cuh:211, :83)."""
    ev = Evaluator(prog, q, mode=mode)
    n = np.asarray(q).shape[0]
    dt = np.float32 if mode == "ref" else np.float64
    return np.stack([np.stack([ev.arg(a, n, dt) for a in s["args"][:3]], -1) for s in spheres], 1)
