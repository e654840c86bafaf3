"""SPICE netlist reader: text -> netlist IR -> ``circuit.Circuit``.

Accepts the reference's subset (parser::parse, src/parser.rs:43-139) with the
same results and the same error texts:

* ``*`` comment lines and blank lines are skipped; ``.END`` stops reading;
* ``R/C/L name n+ n- value``; ``V/I name n+ n- [DC v | v] [AC mag [phase]]
  [PULSE(..) | SIN(..) | PWL(..) | EXP(..)]``; ``D name a k model``;
  ``Q name c b e model``; ``M name d g s [bulk] model`` (bulk ignored);
* ``.MODEL name D|NPN|PNP|NMOS|PMOS(key=value ...)``;
* ``.DC`` / ``.OP`` (operating point), ``.AC DEC|OCT|LIN np fstart fstop``,
  ``.TRAN tstep tstop [tstart] [UIC]``; other dot commands are ignored;
* values take the engineering suffixes T G MEG K M U N P F (any case).

Node names are runs of letters, digits and ``_``; a value is a decimal
number (sign, digits, optional fraction and exponent, or nan / inf /
infinity) followed by an optional suffix, and whatever follows a value on an
element line is ignored, as in the reference (nom 8.0.0's ``double`` and the
parser's tuple grammars, src/parser.rs:154-190).

The netlist is parsed completely before a Circuit is built, so ``.MODEL``
lines may follow the devices that use them (the reference's compiler looks
models up by exact name after parsing, src/compiler.rs:412-424).
"""
from __future__ import annotations

import re
from dataclasses import dataclass, field

from . import circuit as ir
from ._native import OhmnivoreError

F64_MAX = 1.7976931348623157e308
PARSE = 1  # OhmnivoreError code of the Parse variant


def parse_error(msg: str) -> OhmnivoreError:
    """OhmnivoreError::Parse (src/error.rs:5-6)."""
    return OhmnivoreError(PARSE, f"Parse error: {msg}")


class _NoMatch(Exception):
    """A grammar rule did not match (the caller turns it into a message)."""


# ---------------------------------------------------------------------------
# Lexical rules
# ---------------------------------------------------------------------------
_MANTISSA = re.compile(r"[+-]?(?:[0-9]+(?:\.[0-9]*)?|\.[0-9]+)")
_EXPONENT = re.compile(r"[eE][+-]?[0-9]+")
_SPECIALS = ("nan", "infinity", "inf")  # tried in this order, case-insensitive
# (suffix, multiplier), longest first: "MEG" must win over "M"
_SUFFIXES = (("MEG", 1e6), ("T", 1e12), ("G", 1e9), ("K", 1e3), ("M", 1e-3), ("U", 1e-6),
             ("N", 1e-9), ("P", 1e-12), ("F", 1e-15))


def _number(s: str) -> tuple[float, str]:
    """A float at the start of s (nom 8.0.0 number::complete::double): an
    exponent marker must be followed by digits, otherwise nothing matches."""
    m = _MANTISSA.match(s)
    if m:
        end = m.end()
        if end < len(s) and s[end] in "eE":
            e = _EXPONENT.match(s, end)
            if not e:
                raise _NoMatch
            end = e.end()
        return float(s[:end]), s[end:]
    low = s[:8].lower()
    for word in _SPECIALS:
        if low.startswith(word):
            return float(word), s[len(word):]
    raise _NoMatch


def eng_value(s: str) -> tuple[float, str]:
    """Number plus optional engineering suffix (eng_value, src/parser.rs:156-179).
    Returns (value, rest of s)."""
    num, rest = _number(s)
    up = rest[:3].upper()
    for suf, mult in _SUFFIXES:
        if up.startswith(suf):
            return num * mult, rest[len(suf):]
    return num, rest


def parse_value(s: str) -> float:
    """Whole-token convenience form of eng_value (trailing text ignored)."""
    try:
        return eng_value(s.strip())[0]
    except _NoMatch:
        raise ValueError(f"not a SPICE value: {s!r}") from None


def _node_id(s: str) -> tuple[str, str]:
    """Letters, digits and '_' (node_id, src/parser.rs:186-188)."""
    k = 0
    while k < len(s) and (s[k].isalnum() or s[k] == "_"):
        k += 1
    if k == 0:
        raise _NoMatch
    return s[:k], s[k:]


def _element_name(s: str) -> tuple[str, str]:
    """Everything up to the first whitespace (src/parser.rs:191-193)."""
    k = 0
    while k < len(s) and not s[k].isspace():
        k += 1
    if k == 0:
        raise _NoMatch
    return s[:k], s[k:]


def _space1(s: str) -> str:
    """One or more spaces or tabs."""
    k = 0
    while k < len(s) and s[k] in " \t":
        k += 1
    if k == 0:
        raise _NoMatch
    return s[k:]


def _space0(s: str) -> str:
    return s.lstrip(" \t")


def _eng_values(inner: str) -> list[float]:
    """Whitespace-separated values (parse_eng_values, src/parser.rs:442-453)."""
    out = []
    rest = inner.strip()
    while rest:
        try:
            v, r = eng_value(rest)
        except _NoMatch:
            raise ValueError(f"failed to parse value in '{rest}'") from None
        out.append(v)
        rest = r.lstrip()
    return out


# ---------------------------------------------------------------------------
# Netlist IR (src/ir.rs:11-178)
# ---------------------------------------------------------------------------
@dataclass
class Element:
    """One component line. ``kind`` is the element letter R C L V I D Q M;
    ``nodes`` holds 2 names (R C L V I D) or 3 (Q: c b e; M: d g s)."""
    kind: str
    name: str
    nodes: tuple
    value: float | None = None                   # R, C, L
    dc: float | None = None                      # V, I
    ac: tuple[float, float] | None = None        # V, I: (magnitude, phase in degrees)
    tran: object | None = None                   # V, I: circuit.Pulse/Sin/Exp/Pwl
    model: str | None = None                     # D, Q, M


@dataclass
class DiodeModel:
    name: str
    is_: float = 1e-14
    n: float = 1.0


@dataclass
class BjtModel:
    name: str
    is_: float = 1e-16
    bf: float = 100.0
    br: float = 1.0
    nf: float = 1.0
    nr: float = 1.0
    is_npn: bool = True


@dataclass
class MosfetModel:
    name: str
    vto: float = 1.0
    kp: float = 2e-5
    lambda_: float = 0.0
    is_nmos: bool = True


@dataclass
class Dc:
    """.DC / .OP: DC operating point."""


@dataclass
class Ac:
    sweep: str  # "dec" | "oct" | "lin"
    n_points: int
    f_start: float
    f_stop: float


@dataclass
class Tran:
    tstep: float
    tstop: float
    tstart: float = 0.0
    uic: bool = False


@dataclass
class Netlist:
    """Parsed netlist in line order (ir::Circuit, src/ir.rs:171-178)."""
    components: list = field(default_factory=list)
    analyses: list = field(default_factory=list)
    models: list = field(default_factory=list)        # diode models
    bjt_models: list = field(default_factory=list)
    mosfet_models: list = field(default_factory=list)

    def to_circuit(self) -> ir.Circuit:
        """Circuit builder input for compile(): models first (so lookups see
        every .MODEL line), then components in netlist order, which fixes
        node order and therefore the MNA pattern."""
        ckt = ir.Circuit()
        for m in self.models:
            ckt.add_diode_model(m.name, m.is_, m.n)
        for m in self.bjt_models:
            ckt.add_bjt_model(m.name, m.is_, m.bf, m.br, m.nf, m.nr, m.is_npn)
        for m in self.mosfet_models:
            ckt.add_mosfet_model(m.name, m.vto, m.kp, m.lambda_, m.is_nmos)
        for e in self.components:
            a, b = e.nodes[0], e.nodes[1]
            if e.kind == "R":
                ckt.resistor(e.name, a, b, e.value)
            elif e.kind == "C":
                ckt.capacitor(e.name, a, b, e.value)
            elif e.kind == "L":
                ckt.inductor(e.name, a, b, e.value)
            elif e.kind == "V":
                ckt.vsource(e.name, a, b, dc=e.dc, ac=e.ac, tran=e.tran)
            elif e.kind == "I":
                ckt.isource(e.name, a, b, dc=e.dc, ac=e.ac, tran=e.tran)
            elif e.kind == "D":
                ckt.diode(e.name, a, b, e.model)
            elif e.kind == "Q":
                ckt.bjt(e.name, a, b, e.nodes[2], e.model)
            else:
                ckt.mosfet(e.name, a, b, e.nodes[2], e.model)
        return ckt


# ---------------------------------------------------------------------------
# Line rules; each raises ValueError(detail) on a malformed line
# ---------------------------------------------------------------------------
def _rlc(line: str, kind: str) -> Element:
    """Rname n+ n- value (parse_rlc_line, src/parser.rs:200-224)."""
    try:
        name, r = _element_name(line)
        a, r = _node_id(_space1(r))
        b, r = _node_id(_space1(r))
        value, _ = eng_value(_space1(r))
    except _NoMatch:
        raise ValueError(f"failed to parse {kind} component") from None
    return Element(kind, name, (a, b), value=value)


_TRAN_WORDS = ("PULSE", "SIN", "PWL", "EXP")


def _parens(s: str, word: str) -> str:
    """Text between the first '(' and the first ')' (src/parser.rs:376-385)."""
    i = s.find("(")
    if i < 0:
        raise ValueError(f"expected '(' after {word}")
    j = s.find(")")
    if j < i:  # absent, or before the '(' (the reference's slice would panic there)
        raise ValueError(f"expected ')' after {word}")
    return s[i + 1:j]


def _pulse(inner: str) -> ir.Pulse:
    v = _eng_values(inner)
    if len(v) < 2:
        raise ValueError("PULSE requires at least v1 and v2")
    get = lambda k, d: v[k] if k < len(v) else d  # noqa: E731
    return ir.Pulse(v[0], v[1], get(2, 0.0), get(3, 0.0), get(4, 0.0), get(5, F64_MAX),
                    get(6, F64_MAX))


def _sin(inner: str) -> ir.Sin:
    v = _eng_values(inner)
    if len(v) < 3:
        raise ValueError("SIN requires at least vo, va, and freq")
    get = lambda k, d: v[k] if k < len(v) else d  # noqa: E731
    return ir.Sin(v[0], v[1], v[2], get(3, 0.0), get(4, 0.0))


def _pwl(inner: str) -> ir.Pwl:
    v = _eng_values(inner)
    if len(v) < 2 or len(v) % 2:
        raise ValueError("PWL requires an even number of time-value pairs")
    return ir.Pwl([(v[k], v[k + 1]) for k in range(0, len(v), 2)])


def _exp(inner: str) -> ir.Exp:
    v = _eng_values(inner)
    if len(v) < 2:
        raise ValueError("EXP requires at least v1 and v2")
    get = lambda k, d: v[k] if k < len(v) else d  # noqa: E731
    return ir.Exp(v[0], v[1], get(2, 0.0), get(3, F64_MAX), get(4, F64_MAX), get(5, F64_MAX))


_TRAN_RULES = {"PULSE": _pulse, "SIN": _sin, "PWL": _pwl, "EXP": _exp}


def _source_specs(rest: str):
    """[DC v | v] [AC mag [phase]] [waveform] (parse_source_specs,
    src/parser.rs:279-373). A bare leading number is the DC value; text
    that matches none of the forms is ignored."""
    dc = ac = tran = None
    if not rest:
        return dc, ac, tran
    up = rest.upper()
    if up.startswith("DC"):
        rest = rest[2:].lstrip()
        if rest and not rest.upper().startswith("AC"):
            try:
                dc, rest = eng_value(rest)
            except _NoMatch:
                raise ValueError("failed to parse DC value") from None
            rest = rest.lstrip()
    elif not up.startswith(("AC",) + _TRAN_WORDS):
        try:
            dc, r = eng_value(rest)
            rest = r.lstrip()
        except _NoMatch:
            pass
    if rest.upper().startswith("AC"):
        rest = rest[2:].lstrip()
        try:
            mag, rest = eng_value(rest)
        except _NoMatch:
            raise ValueError("failed to parse AC magnitude") from None
        rest = rest.lstrip()
        phase = 0.0
        if rest and not rest.upper().startswith(_TRAN_WORDS):
            try:
                phase, rest = eng_value(rest)
            except _NoMatch:
                raise ValueError("failed to parse AC phase") from None
            rest = rest.lstrip()
        ac = (mag, phase)
    up = rest.upper()
    for word in _TRAN_WORDS:
        if up.startswith(word):
            tran = _TRAN_RULES[word](_parens(rest, word))
            break
    return dc, ac, tran


def _source(line: str, kind: str) -> Element:
    """Vname/Iname n+ n- [specs] (parse_source_line, src/parser.rs:230-268)."""
    try:
        name, r = _element_name(line)
        a, r = _node_id(_space1(r))
        b, r = _node_id(_space1(r))
        r = _space0(r)
    except _NoMatch:
        raise ValueError("failed to parse source name/nodes") from None
    dc, ac, tran = _source_specs(r.strip())
    return Element(kind, name, (a, b), dc=dc, ac=ac, tran=tran)


def _diode(line: str) -> Element:
    """Dname anode cathode MODEL (src/parser.rs:461-480)."""
    try:
        name, r = _element_name(line)
        a, r = _node_id(_space1(r))
        k, r = _node_id(_space1(r))
        model, _ = _node_id(_space1(r))
    except _NoMatch:
        raise ValueError("failed to parse diode element") from None
    return Element("D", name, (a, k), model=model)


def _bjt(line: str) -> Element:
    """Qname collector base emitter MODEL (src/parser.rs:487-508)."""
    try:
        name, r = _element_name(line)
        c, r = _node_id(_space1(r))
        b, r = _node_id(_space1(r))
        e, r = _node_id(_space1(r))
        model, _ = _node_id(_space1(r))
    except _NoMatch:
        raise ValueError("failed to parse BJT element") from None
    return Element("Q", name, (c, b, e), model=model)


def _mosfet(line: str) -> Element:
    """Mname drain gate source [bulk] MODEL (src/parser.rs:515-548)."""
    try:
        name, r = _element_name(line)
        d, r = _node_id(_space1(r))
        g, r = _node_id(_space1(r))
        s, r = _node_id(_space1(r))
        r = _space0(r)
    except _NoMatch:
        raise ValueError("failed to parse MOSFET element") from None
    tokens = r.split()
    if len(tokens) not in (1, 2):
        raise ValueError("expected model name (and optional bulk node) after MOSFET terminals")
    return Element("M", name, (d, g, s), model=tokens[-1])


# key -> attribute, per model family (src/parser.rs:645-734); the first
# matching prefix wins, checked in this order
_DIODE_KEYS = (("IS=", "is_"), ("N=", "n"))
_BJT_KEYS = (("IS=", "is_"), ("BF=", "bf"), ("BR=", "br"), ("NF=", "nf"), ("NR=", "nr"))
_MOS_KEYS = (("VTO=", "vto"), ("KP=", "kp"), ("LAMBDA=", "lambda_"))


def _model_params(params: str, keys, obj, family: str) -> None:
    for token in params.split():
        up = token.upper()
        for key, attr in keys:
            if up.startswith(key):
                text = token[len(key):]
                try:
                    val, _ = eng_value(text)
                except _NoMatch:
                    raise ValueError(f"failed to parse {key[:-1]} value: {text}") from None
                setattr(obj, attr, val)
                break
        else:
            raise ValueError(f"unknown {family} model parameter: {token}")


def _model(line: str):
    """.MODEL name TYPE[(params)] (parse_model_command, src/parser.rs:561-640)."""
    rest = line.strip()
    if not rest.upper().startswith(".MODEL"):
        raise ValueError("expected .MODEL directive")
    rest = rest[6:].lstrip()
    try:
        name, rest = _node_id(rest)
    except _NoMatch:
        raise ValueError("expected model name after .MODEL") from None
    rest = rest.lstrip()
    try:
        mtype, rest = _node_id(rest)
    except _NoMatch:
        raise ValueError("expected model type after model name") from None
    rest = rest.lstrip()
    params = None
    if rest.startswith("("):
        end = rest.find(")")
        if end < 0:
            raise ValueError("missing closing ')' in .MODEL parameters")
        params = rest[1:end]
    t = mtype.upper()
    if t == "D":
        m, keys, fam = DiodeModel(name), _DIODE_KEYS, "diode"
    elif t in ("NPN", "PNP"):
        m, keys, fam = BjtModel(name, is_npn=(t == "NPN")), _BJT_KEYS, "BJT"
    elif t in ("NMOS", "PMOS"):
        m, keys, fam = MosfetModel(name, is_nmos=(t == "NMOS")), _MOS_KEYS, "MOSFET"
    else:
        raise ValueError(f"unsupported model type '{mtype}', expected D, NPN, PNP, NMOS, or PMOS")
    if params is not None:
        _model_params(params, keys, m, fam)
    return m


def _usize(s: str) -> tuple[int, str]:
    k = 0
    while k < len(s) and s[k] in "0123456789":
        k += 1
    if k == 0:
        raise ValueError("expected integer")
    n = int(s[:k])
    if n >= 1 << 64:
        raise ValueError("invalid integer")
    return n, s[k:]


def _ac(line: str) -> Ac:
    """.AC DEC|OCT|LIN np fstart fstop (parse_ac_command, src/parser.rs:765-822)."""
    rest = line[3:].lstrip()
    sweep = rest[:3].lower()
    if sweep not in ("dec", "oct", "lin"):
        raise ValueError("expected DEC, OCT, or LIN for AC sweep type")
    rest = rest[3:].lstrip()
    n, rest = _usize(rest)
    rest = rest.lstrip()
    try:
        f0, rest = eng_value(rest)
    except _NoMatch:
        raise ValueError("failed to parse AC f_start") from None
    rest = rest.lstrip()
    try:
        f1, _ = eng_value(rest)
    except _NoMatch:
        raise ValueError("failed to parse AC f_stop") from None
    return Ac(sweep, n, f0, f1)


def _tran(line: str) -> Tran:
    """.TRAN tstep tstop [tstart] [UIC] (parse_tran_command, src/parser.rs:825-858).
    A value after tstart (a maximum step) is ignored, and so is UIC after it."""
    rest = line[5:].lstrip()
    try:
        tstep, rest = eng_value(rest)
    except _NoMatch:
        raise ValueError("failed to parse TRAN tstep") from None
    rest = rest.lstrip()
    try:
        tstop, rest = eng_value(rest)
    except _NoMatch:
        raise ValueError("failed to parse TRAN tstop") from None
    rest = rest.lstrip()
    tstart, uic = 0.0, False
    if rest:
        if rest.upper().startswith("UIC"):
            uic = True
        else:
            try:
                tstart, r = eng_value(rest)
                uic = r.lstrip().upper().startswith("UIC")
            except _NoMatch:
                pass
    return Tran(tstep, tstop, tstart, uic)


def _dot(line: str):
    """.DC/.OP, .AC, .TRAN; None for other dot commands (src/parser.rs:743-762).
    Only a bare .DC requests the operating point: .DC with sweep arguments
    is one of the ignored commands."""
    t = line.upper().strip()
    if t in (".DC", ".OP"):
        return Dc()
    if t.startswith(".AC"):
        return _ac(line)
    if t.startswith(".TRAN"):
        return _tran(line)
    return None


_RULES = {"R": lambda s: _rlc(s, "R"), "C": lambda s: _rlc(s, "C"), "L": lambda s: _rlc(s, "L"),
          "V": lambda s: _source(s, "V"), "I": lambda s: _source(s, "I"), "D": _diode,
          "Q": _bjt, "M": _mosfet}


def _lines(text: str):
    """str::lines: split at '\\n', dropping one '\\r' before it."""
    for raw in text.split("\n"):
        yield raw[:-1] if raw.endswith("\r") else raw


def parse(text: str) -> Netlist:
    """parser::parse (src/parser.rs:43-139). Raises OhmnivoreError (kind
    "Parse") naming the 1-based line and quoting it."""
    out = Netlist()
    for num, raw in enumerate(_lines(text), start=1):
        line = raw.strip()
        if not line or line.startswith("*"):
            continue
        up = line.upper()
        if up == ".END":
            break
        first = line[0]
        try:
            if first.upper() in _RULES:
                out.components.append(_RULES[first.upper()](line))
            elif first == ".":
                if up.startswith(".MODEL"):
                    m = _model(line)
                    (out.models if isinstance(m, DiodeModel) else
                     out.bjt_models if isinstance(m, BjtModel) else out.mosfet_models).append(m)
                else:
                    a = _dot(line)
                    if a is not None:
                        out.analyses.append(a)
            else:
                raise parse_error(f"line {num}: unknown element '{first}': {raw}")
        except ValueError as e:
            raise parse_error(f"line {num}: {e} in: {raw}") from None
    return out


def parse_file(path: str) -> Netlist:
    with open(path, encoding="utf-8") as f:
        return parse(f.read())
