"""VACASK netlist front end: ``.sim`` text -> parsed circuit -> flattened
leaf instances -> :class:`~vajax_amd.netlist.Netlist`.

Follows the reference's parser and flattening (vajax/netlist/parser.py:42-762,
vajax/netlist/circuit.py, vajax/analysis/parsing.py:32-230,
vajax/analysis/engine.py:256-263, 386-403, 541-569,
vajax/utils/safe_eval.py:19-200):

* statements: a title line, ``load``, ``include`` (read relative to the
  including file), ``global``, ``ground``, ``model name module [(] p=v ... [)]``,
  ``parameters``, ``subckt name(ports) ... ends``, instances
  ``name (n1 n2 ...) master p=v ...`` (terminal lists and model parameter
  lists may span lines), ``control ... endc`` (options, analysis, save, var,
  print; the others skipped), ``embed`` blocks and ``@if`` conditionals
  skipped as the reference skips them;
* flattening: subcircuit instances expand recursively, inner names are
  ``<instance path>.<name>``, ports map to the caller's nodes, globals and the
  ground keep their names; a subcircuit's ``parameters`` defaults are
  evaluated in the caller's environment extended in order, then the
  instance's parameters, evaluated in the caller's environment, override them;
* expressions: numbers with SPICE scale suffixes, parameter names, ``+ - * /
  **`` and unary signs, and a whitelist of math functions -- evaluated over
  Python's ``ast`` without ``eval``;
* ``generate`` blocks (netlist-embedded scripts that the reference runs to
  produce netlist text) are never executed: the array multiplier block of the
  reference's mul64 benchmark is replaced by the same topology from an
  in-tree generator, any other block is refused.

Device mapping onto the models built here: ``resistor``/``sp_resistor``,
``capacitor``/``sp_capacitor``, ``vsource``, ``isource``, ``diode``/``sp_diode``
(junction diode: ``is``, ``n``, ``cj`` or ``cjo``), ``ekv`` (the EKV
long-channel module, parameters by name), ``ekv_va`` (the EKV 2.6 module of
vajax/devices/models/ekv/ekv.va, parameters by name) and ``psp103va`` (no PSP103 here: its
instances become the EKV substitute of the synthetic circuits, with the
card's ``type`` polarity and the instance's ``w``/``l``).
"""

from __future__ import annotations

import ast
import functools
import math
import operator
import re
from dataclasses import dataclass, field
from pathlib import Path
from typing import Callable, Dict, List, Optional, Tuple, Union

from .generators import NMOS_DEFAULTS, PMOS_DEFAULTS
from .netlist import MODEL_PARAMS, SOURCE_DEFAULTS, Netlist, NetlistBuilder

# ---------------------------------------------------------------------------
# Numbers and expressions
# ---------------------------------------------------------------------------

# SPICE scale suffixes, longest first; letters after a scale (units) are ignored
_SCALES: Tuple[Tuple[str, float], ...] = (
    ("meg", 1e6), ("mil", 25.4e-6), ("t", 1e12), ("g", 1e9), ("k", 1e3), ("m", 1e-3),
    ("u", 1e-6), ("n", 1e-9), ("p", 1e-12), ("f", 1e-15), ("a", 1e-18),
)
_NUM = re.compile(r"(?<![A-Za-z_0-9.])((?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?)([A-Za-z]*)")
_FUNCS: Dict[str, Callable] = {
    "sin": math.sin, "cos": math.cos, "tan": math.tan, "asin": math.asin, "acos": math.acos,
    "atan": math.atan, "atan2": math.atan2, "sinh": math.sinh, "cosh": math.cosh,
    "tanh": math.tanh, "exp": math.exp, "log": math.log, "ln": math.log, "log10": math.log10,
    "sqrt": math.sqrt, "abs": abs, "pow": pow, "min": min, "max": max, "floor": math.floor,
    "ceil": math.ceil,
}
_CONSTS = {"pi": math.pi}
_BINOPS = {ast.Add: operator.add, ast.Sub: operator.sub, ast.Mult: operator.mul,
           ast.Div: operator.truediv, ast.Pow: operator.pow, ast.Mod: operator.mod}
_UNOPS = {ast.UAdd: operator.pos, ast.USub: operator.neg}


def _scale(letters: str) -> float:
    low = letters.lower()
    for s, mult in _SCALES:
        if low.startswith(s):
            return mult
    return 1.0


@functools.lru_cache(maxsize=65536)
def parse_number(text: str) -> Optional[float]:
    """A number with an optional scale suffix ("1u", "0.1n", "1.5meg", "10pF",
    "2e-3"); None when ``text`` is not one."""
    t = text.strip().strip('"')
    m = _NUM.fullmatch(t)
    if not m:
        return None
    return float(m.group(1)) * _scale(m.group(2))


class ExpressionError(ValueError):
    pass


def evaluate(expr: Union[str, float], env: Dict[str, float]) -> float:
    """Evaluate a parameter expression with names from ``env``."""
    if not isinstance(expr, str):
        return float(expr)
    v = parse_number(expr)
    if v is not None:
        return v
    return float(_eval_node(_parse_expr(expr), env, expr))


@functools.lru_cache(maxsize=65536)
def _parse_expr(expr: str):
    """The syntax tree of an expression (scale suffixes folded into numbers)."""
    src = _NUM.sub(lambda m: repr(float(m.group(1)) * _scale(m.group(2))), expr.strip())
    try:
        return ast.parse(src, mode="eval").body
    except SyntaxError as e:
        raise ExpressionError(f"bad expression {expr!r}") from e


def _eval_node(n, env, expr):
    if isinstance(n, ast.Constant) and isinstance(n.value, (int, float)):
        return n.value
    if isinstance(n, ast.Name):
        if n.id in env:
            return env[n.id]
        if n.id in _CONSTS:
            return _CONSTS[n.id]
        raise ExpressionError(f"unknown name {n.id!r} in {expr!r}")
    if isinstance(n, ast.BinOp) and type(n.op) in _BINOPS:
        return _BINOPS[type(n.op)](_eval_node(n.left, env, expr), _eval_node(n.right, env, expr))
    if isinstance(n, ast.UnaryOp) and type(n.op) in _UNOPS:
        return _UNOPS[type(n.op)](_eval_node(n.operand, env, expr))
    if (isinstance(n, ast.Call) and isinstance(n.func, ast.Name) and n.func.id in _FUNCS
            and not n.keywords):
        return _FUNCS[n.func.id](*[_eval_node(a, env, expr) for a in n.args])
    raise ExpressionError(f"unsupported construct in {expr!r}")


# ---------------------------------------------------------------------------
# Lexer
# ---------------------------------------------------------------------------

KEYWORDS = {"load", "include", "global", "ground", "model", "parameters", "subckt", "ends",
            "control", "endc", "embed", "options", "analysis", "save", "var", "print", "alter",
            "postprocess", "abort", "clear", "elaborate"}


@dataclass
class Tok:
    kind: str  # NL NAME STRING VALUE LP RP EQ DIR KW EOF
    text: str
    line: int


def tokenize(text: str) -> List[Tok]:
    """Tokens of a netlist.  The text after ``=`` up to whitespace, a comma or
    an unbalanced closing bracket is one VALUE token (expressions may hold
    parentheses); ``//`` and ``/* */`` are comments."""
    out: List[Tok] = []
    i, n, line = 0, len(text), 1
    while i < n:
        c = text[i]
        if c in " \t\r":
            i += 1
        elif c == "\n":
            out.append(Tok("NL", "\n", line))
            line += 1
            i += 1
        elif text.startswith("//", i):
            while i < n and text[i] != "\n":
                i += 1
        elif text.startswith("/*", i):
            j = text.find("*/", i + 2)
            j = n if j < 0 else j + 2
            line += text.count("\n", i, j)
            i = j
        elif c == "@":
            j = i + 1
            while j < n and text[j].isalpha():
                j += 1
            word = text[i:j]
            kind = "DIR" if word.lower() in ("@if", "@elseif", "@else", "@endif", "@end") else "NAME"
            out.append(Tok(kind, word, line))
            i = j
        elif c == '"':
            j = text.find('"', i + 1)
            j = n if j < 0 else j + 1
            out.append(Tok("STRING", text[i:j], line))
            i = j
        elif c == "(":
            out.append(Tok("LP", c, line))
            i += 1
        elif c == ")":
            out.append(Tok("RP", c, line))
            i += 1
        elif c == "=":
            out.append(Tok("EQ", c, line))
            i += 1
            while i < n and text[i] in " \t":
                i += 1
            if i < n and text[i] == '"':
                continue
            j, depth = i, 0
            while j < n:
                ch = text[j]
                if ch in "([":
                    depth += 1
                elif ch in ")]":
                    if depth == 0:
                        break
                    depth -= 1
                elif ch in " \t\r\n," and depth == 0:
                    break
                j += 1
            if j > i:
                out.append(Tok("VALUE", text[i:j], line))
            i = j
        elif c.isalnum() or c == "_":
            j = i
            while j < n and (text[j].isalnum() or text[j] in "_."):
                j += 1
            word = text[i:j]
            out.append(Tok("KW" if word.lower() in KEYWORDS else "NAME", word, line))
            i = j
        else:
            i += 1  # other punctuation (commas, ``<<<`` markers) separates tokens
    out.append(Tok("EOF", "", line))
    return out


# ---------------------------------------------------------------------------
# Parsed circuit (vajax/netlist/circuit.py)
# ---------------------------------------------------------------------------

@dataclass
class Instance:
    name: str
    terminals: List[str]
    model: str
    params: Dict[str, str] = field(default_factory=dict)


@dataclass
class Model:
    name: str
    module: str
    params: Dict[str, str] = field(default_factory=dict)


@dataclass
class Subcircuit:
    name: str
    terminals: List[str]
    params: Dict[str, str] = field(default_factory=dict)
    instances: List[Instance] = field(default_factory=list)
    models: Dict[str, Model] = field(default_factory=dict)


@dataclass
class Analysis:
    name: str
    kind: str
    params: Dict[str, str] = field(default_factory=dict)


@dataclass
class Control:
    options: Dict[str, str] = field(default_factory=dict)
    analyses: List[Analysis] = field(default_factory=list)
    saves: List[str] = field(default_factory=list)
    vars: Dict[str, str] = field(default_factory=dict)
    prints: List[Tuple[str, List[str]]] = field(default_factory=list)
    elaborate: Optional[str] = None


@dataclass
class Circuit:
    title: str = ""
    ground: str = "0"
    globals: List[str] = field(default_factory=list)
    loads: List[str] = field(default_factory=list)
    includes: List[str] = field(default_factory=list)
    params: Dict[str, str] = field(default_factory=dict)
    models: Dict[str, Model] = field(default_factory=dict)
    subckts: Dict[str, Subcircuit] = field(default_factory=dict)
    top_instances: List[Instance] = field(default_factory=list)
    control: Optional[Control] = None


class NetlistSyntaxError(SyntaxError):
    pass


class _Parser:
    def __init__(self, toks: List[Tok], circuit: Circuit, base: Optional[Path],
                 depth: int = 0):
        self.t = toks
        self.i = 0
        self.c = circuit
        self.base = base
        self.depth = depth
        self.sub: Optional[Subcircuit] = None

    # -- token helpers --
    def cur(self, k: int = 0) -> Tok:
        return self.t[min(self.i + k, len(self.t) - 1)]

    def take(self) -> Tok:
        tok = self.cur()
        self.i += 1
        return tok

    def want(self, kind: str, word: Optional[str] = None) -> Tok:
        tok = self.cur()
        if tok.kind != kind or (word is not None and tok.text.lower() != word):
            raise NetlistSyntaxError(f"line {tok.line}: expected {word or kind}, got {tok.text!r}")
        return self.take()

    def is_kw(self, word: str) -> bool:
        return self.cur().kind == "KW" and self.cur().text.lower() == word

    def skip_nl(self):
        while self.cur().kind == "NL":
            self.i += 1

    def rest_of_line(self):
        while self.cur().kind not in ("NL", "EOF"):
            self.i += 1

    # -- grammar --
    def parse(self, top: bool):
        self.skip_nl()
        if top and self.cur().kind == "NAME" and self.cur(1).kind not in ("LP", "EQ"):
            words = []
            while self.cur().kind in ("NAME", "STRING", "KW"):
                words.append(self.take().text)
            self.c.title = " ".join(words)
        while self.cur().kind != "EOF":
            self.statement()
            self.skip_nl()

    def statement(self):
        tok = self.cur()
        word = tok.text.lower()
        if tok.kind == "NL":
            self.i += 1
        elif tok.kind == "DIR":
            self.conditional()
        elif tok.kind == "NAME" and word == "generate":
            raise NetlistSyntaxError(
                f"line {tok.line}: generate blocks run netlist-embedded scripts; "
                "expand them to netlist text first")
        elif tok.kind == "NAME":
            self.instance()
        elif tok.kind == "KW" and word == "load":
            self.i += 1
            self.c.loads.append(self.want("STRING").text[1:-1])
        elif tok.kind == "KW" and word == "include":
            self.i += 1
            self.include(self.want("STRING").text[1:-1])
        elif tok.kind == "KW" and word == "global":
            self.i += 1
            while self.cur().kind == "NAME":
                self.c.globals.append(self.take().text)
        elif tok.kind == "KW" and word == "ground":
            self.i += 1
            self.c.ground = self.want("NAME").text
        elif tok.kind == "KW" and word == "model":
            self.model()
        elif tok.kind == "KW" and word == "parameters":
            self.i += 1
            (self.sub.params if self.sub else self.c.params).update(self.params())
        elif tok.kind == "KW" and word == "subckt":
            self.subckt()
        elif tok.kind == "KW" and word == "control":
            self.control()
        elif tok.kind == "KW" and word == "embed":
            while self.cur().kind != "EOF" and ">>>FILE" not in self.cur().text:
                self.i += 1
            self.i += 1
        else:
            raise NetlistSyntaxError(f"line {tok.line}: unexpected {tok.text!r}")

    def include(self, name: str):
        self.c.includes.append(name)
        if self.base is None:
            return
        path = (self.base / name) if not Path(name).is_absolute() else Path(name)
        if not path.exists():
            raise FileNotFoundError(f"include {name!r}: {path} not found")
        if self.depth > 16:
            raise NetlistSyntaxError(f"include {name!r}: nested too deeply")
        inner = _Parser(tokenize(expand_generate_blocks(path.read_text())), self.c, path.parent,
                        self.depth + 1)
        inner.parse(top=False)

    def params(self, multiline: bool = False) -> Dict[str, str]:
        out: Dict[str, str] = {}
        while True:
            if multiline:
                self.skip_nl()
            if not (self.cur().kind in ("NAME", "KW") and self.cur(1).kind == "EQ"):
                return out
            name = self.take().text
            self.want("EQ")
            val = self.cur()
            if val.kind not in ("STRING", "NAME", "VALUE"):
                raise NetlistSyntaxError(f"line {val.line}: expected a value for {name}")
            out[name] = self.take().text

    def model(self):
        self.want("KW", "model")
        name = self.want("NAME").text
        module = self.want("NAME").text
        if self.cur().kind == "LP":
            self.i += 1
            params = self.params(multiline=True)
            self.skip_nl()
            self.want("RP")
        else:
            params = self.params()
        m = Model(name, module, params)
        if self.sub is not None:
            self.sub.models[name] = m
        else:
            self.c.models[name] = m

    def subckt(self):
        self.want("KW", "subckt")
        name = self.want("NAME").text
        self.want("LP")
        ports = []
        while self.cur().kind in ("NAME", "NL"):
            tok = self.take()
            if tok.kind == "NAME":
                ports.append(tok.text)
        self.want("RP")
        if self.sub is not None:
            raise NetlistSyntaxError(f"line {self.cur().line}: nested subckt {name!r}")
        self.sub = Subcircuit(name, ports)
        self.skip_nl()
        while not self.is_kw("ends"):
            if self.cur().kind == "EOF":
                raise NetlistSyntaxError(f"subckt {name!r}: missing ends")
            self.statement()
            self.skip_nl()
        self.i += 1
        self.c.subckts[name] = self.sub
        self.sub = None

    def instance(self):
        name = self.want("NAME").text
        self.want("LP")
        terms = []
        while self.cur().kind in ("NAME", "NL"):
            tok = self.take()
            if tok.kind == "NAME":
                terms.append(tok.text)
        self.want("RP")
        master = self.want("NAME").text
        inst = Instance(name, terms, master, self.params())
        (self.sub.instances if self.sub else self.c.top_instances).append(inst)

    def conditional(self):
        # @if ... @endif blocks are skipped whole, as the reference skips them
        word = self.take().text.lower()
        self.rest_of_line()
        if word != "@if":
            return
        depth = 1
        while depth and self.cur().kind != "EOF":
            tok = self.take()
            if tok.kind == "DIR":
                w = tok.text.lower()
                depth += 1 if w == "@if" else -1 if w in ("@endif", "@end") else 0

    def control(self):
        self.want("KW", "control")
        ctl = self.c.control or Control()
        self.c.control = ctl
        while True:
            self.skip_nl()
            tok = self.cur()
            if tok.kind == "EOF":
                raise NetlistSyntaxError("control block: missing endc")
            word = tok.text.lower()
            if tok.kind == "KW" and word == "endc":
                self.i += 1
                return
            self.i += 1
            if tok.kind == "KW" and word == "options":
                ctl.options.update(self.params())
            elif tok.kind == "KW" and word == "analysis":
                name = self.want("NAME").text
                kind = self.want("NAME").text
                ctl.analyses.append(Analysis(name, kind, self.params()))
            elif tok.kind == "KW" and word == "var":
                ctl.vars.update(self.params())
            elif tok.kind == "KW" and word == "save":
                while self.cur().kind == "NAME":
                    sig = self.take().text
                    if self.cur().kind == "LP":
                        self.i += 1
                        args = []
                        while self.cur().kind == "NAME":
                            args.append(self.take().text)
                        self.want("RP")
                        sig = f"{sig}({','.join(args)})"
                    ctl.saves.append(sig)
            elif tok.kind == "KW" and word == "print":
                sub = self.take().text if self.cur().kind == "NAME" else ""
                args = []
                while self.cur().kind in ("NAME", "STRING"):
                    args.append(self.take().text.strip('"'))
                ctl.prints.append((sub, args))
            elif tok.kind == "KW" and word == "elaborate":
                # elaborate circuit("name") (parsing.py:17-29)
                if self.cur().kind == "NAME" and self.cur(1).kind == "LP":
                    self.i += 2
                    if self.cur().kind == "STRING":
                        ctl.elaborate = self.take().text[1:-1]
                self.rest_of_line()
            else:
                self.rest_of_line()  # postprocess, alter, abort, clear, unknown


_GENERATE = re.compile(r'^[ \t]*generate[ \t]+"([^"]+)"[ \t]*<<<FILE[ \t]*\n(.*?)^>>>FILE[ \t]*$',
                       re.M | re.S)
_MULTIPLIER = re.compile(r"multiplier_(\d+)x(\d+)\.py$")


def expand_generate_blocks(text: str) -> str:
    """``generate "<script>" <<<FILE ... >>>FILE`` blocks are scripts the
    reference runs at parse time, splicing their output into the netlist.
    They are never executed here: a block named ``multiplier_<N>x<N>.py``
    (the array multiplier of vajax/benchmarks/data/mul64/runme.sim:8-13) is
    replaced by the same topology from the in-tree generator
    (:func:`vajax_amd.generators.array_multiplier_netlist`); any other block is
    refused."""
    from .generators import array_multiplier_netlist

    def repl(m):
        name = m.group(1)
        mm = _MULTIPLIER.search(name)
        if mm and mm.group(1) == mm.group(2):
            return (f"// generate {name}: array multiplier from the in-tree generator\n"
                    + array_multiplier_netlist(int(mm.group(1))) + "\n")
        raise NetlistSyntaxError(f"generate {name!r}: netlist-embedded scripts are not "
                                 "executed; expand the block to netlist text first")

    return _GENERATE.sub(repl, text)


def parse_netlist(source: Union[str, Path], text: Optional[str] = None) -> Circuit:
    """Parse a VACASK netlist from a file path, or from ``text`` (includes then
    resolve against ``source`` when it names a directory or file)."""
    base: Optional[Path] = None
    if text is None:
        path = Path(source)
        text = path.read_text()
        base = path.parent
    elif source:
        p = Path(source)
        base = p if p.is_dir() else p.parent
    c = Circuit()
    _Parser(tokenize(expand_generate_blocks(text)), c, base).parse(top=True)
    return c


# ---------------------------------------------------------------------------
# Flattening (vajax/analysis/parsing.py:32-140)
# ---------------------------------------------------------------------------

STRING_PARAMS = {"type", "wave"}


def _value(key: str, expr: str, env: Dict[str, float]):
    s = expr.strip()
    if len(s) >= 2 and s[0] == s[-1] and s[0] in "\"'":
        return s[1:-1]
    if key.lower() in STRING_PARAMS:
        return s
    return evaluate(s, env)


def flatten(c: Circuit) -> List[Tuple[str, List[str], str, Dict[str, object], Dict[str, Model]]]:
    """Leaf instances (path name, terminals, master, evaluated parameters,
    models visible at the instance)."""
    ground = c.ground or "0"
    glob = set(c.globals) | {ground}
    top_env: Dict[str, float] = {}
    for k, v in c.params.items():
        top_env[k] = evaluate(v, top_env)
    tops = list(c.top_instances)
    if c.control is not None and c.control.elaborate:
        sc = c.subckts.get(c.control.elaborate)
        if sc is None:
            raise KeyError(f"elaborate: no subckt {c.control.elaborate!r}")
        tops.append(Instance("top", list(sc.terminals), sc.name, {}))
    out = []

    def node(t, prefix, ports):
        if t in ports:
            return ports[t]
        if t in glob or not prefix:
            return t
        return f"{prefix}.{t}"

    def walk(inst: Instance, prefix: str, ports: Dict[str, str], env: Dict[str, float],
             models: Dict[str, Model], depth: int):
        if depth > 64:
            raise RecursionError(f"subckt nesting too deep at {inst.name!r}")
        sc = c.subckts.get(inst.model)
        path = f"{prefix}.{inst.name}" if prefix else inst.name
        if sc is None:
            terms = [node(t, prefix, ports) for t in inst.terminals]
            params = {k: _value(k, v, env) for k, v in inst.params.items()}
            out.append((path, terms, inst.model, params, models))
            return
        if len(inst.terminals) != len(sc.terminals):
            raise ValueError(f"{path}: {sc.name} has {len(sc.terminals)} ports, "
                             f"{len(inst.terminals)} given")
        inner_ports = {p: node(t, prefix, ports) for p, t in zip(sc.terminals, inst.terminals)}
        inner_env = dict(env)
        for k, v in sc.params.items():
            inner_env[k] = _value(k, v, inner_env)
        for k, v in inst.params.items():
            inner_env[k] = _value(k, v, env)
        inner_models = dict(models)
        inner_models.update(sc.models)
        for sub in sc.instances:
            walk(sub, path, inner_ports, inner_env, inner_models, depth + 1)

    for inst in tops:
        walk(inst, "", {}, top_env, dict(c.models), 0)
    return out


# ---------------------------------------------------------------------------
# Mapping to the models built here
# ---------------------------------------------------------------------------

MODULE_KIND = {
    "resistor": "resistor", "sp_resistor": "resistor",
    "capacitor": "capacitor", "sp_capacitor": "capacitor",
    "vsource": "vsource", "isource": "isource",
    "diode": "diode", "sp_diode": "diode",
    "ekv": "mosfet", "ekv_longchannel": "mosfet", "psp103va": "mosfet_psp",
    "ekv_va": "ekv26", "ekv26": "ekv26",
}
_ALIASES = {"diode": {"cjo": "cj", "cj0": "cj"}}
# netlist options this build reads (options.py names)
_OPTIONS = {"tran_method", "reltol", "abstol", "vntol", "gmin", "tran_lteratio",
            "tran_redofactor", "tran_fs", "tran_ft", "tran_minpts", "maxstep", "tran_itl",
            "tran_gshunt", "icmode"}


def _num(v) -> float:
    return float(v) if not isinstance(v, str) else evaluate(v, {})


def to_netlist(c: Circuit, title: Optional[str] = None) -> Netlist:
    """Flatten and map onto the device models built here."""
    b = NetlistBuilder(title or c.title or "netlist", ground=c.ground or "0")
    card_values: Dict[int, Dict[str, object]] = {}
    for path, terms, master, params, models in flatten(c):
        model = models.get(master)
        module = (model.module if model else master).lower()
        kind = MODULE_KIND.get(module)
        if kind is None:
            raise NotImplementedError(f"{path}: no device model for module {module!r}")
        mparams: Dict[str, object] = {}
        if model is not None:
            # a model card's parameters are constants: evaluated once per card
            mparams = card_values.get(id(model))
            if mparams is None:
                mparams = {k: _value(k, v, {}) for k, v in model.params.items()}
                card_values[id(model)] = mparams
        p = {k.lower(): v for k, v in {**mparams, **params}.items()}
        if kind in ("vsource", "isource"):
            if len(terms) != 2:
                raise ValueError(f"{path}: a source has 2 terminals")
            wave = str(p.pop("type", "dc")).strip('"').lower()
            points = p.pop("wave", None)
            if isinstance(points, str):
                body = points.strip()
                if not (body.startswith("[") and body.endswith("]")):
                    raise ValueError(f"{path}: pwl wave must be a [t0, v0, t1, v1, ...] list")
                body = body[1:-1].strip()
                points = [evaluate(x, {}) for x in body.split(",")] if body else []
            if "sinephase" in p and "phase" not in p:
                p["phase"] = p.pop("sinephase")
            kw = {k: _num(v) for k, v in p.items() if k in SOURCE_DEFAULTS}
            b.source(kind, path, terms[0], terms[1], type=wave,
                     wave=points if wave == "pwl" else None, **kw)
            continue
        if kind == "mosfet_psp":
            # PSP103 is not available: the EKV substitute of the synthetic
            # circuits (NMOS_DEFAULTS / PMOS_DEFAULTS) with the card's polarity
            # and the instance's size
            kw = dict(NMOS_DEFAULTS if _num(p.get("type", 1)) >= 0 else PMOS_DEFAULTS)
            for k in ("w", "l"):
                if k in p:
                    kw[k] = _num(p[k])
            b.device("mosfet", path, terms, **kw)
            continue
        alias = _ALIASES.get(kind, {})
        known = MODEL_PARAMS[kind]
        kw = {}
        for k, v in p.items():
            k = alias.get(k, k)
            if k in known:
                kw[k] = _num(v)
        if kind == "mosfet" and "type" not in kw:
            kw["type"] = 1.0
        b.device(kind, path, terms, **kw)
    analysis: Dict[str, float] = {}
    options: Dict[str, object] = {}
    ctl = c.control
    if ctl is not None:
        for k, v in ctl.options.items():
            if k.lower() in _OPTIONS:
                s = v.strip('"')
                options[k.lower()] = s if k.lower() in ("tran_method", "icmode") else _num(s)
        for a in ctl.analyses:
            if a.kind.lower() == "tran":
                for k, v in a.params.items():
                    s = v.strip('"')
                    if k.lower() == "icmode":
                        options["icmode"] = s
                    elif k.lower() in ("step", "stop", "maxstep", "start"):
                        analysis[k.lower()] = _num(s)
                break
    return b.build(analysis=analysis, options=options)


def load(source: Union[str, Path], text: Optional[str] = None) -> Netlist:
    """Parse, flatten and map a VACASK netlist."""
    c = parse_netlist(source, text)
    return to_netlist(c)
