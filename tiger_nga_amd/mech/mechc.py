"""Mechanism compiler: reads a Tiger_NGA/FlameMaster-style Fortran mechanism
(``*F.f90`` + ``*F90.h``) as text and lowers it to rate tables.

The mechanism files are the reference's link-time chemistry plugin (the
11-routine ABI of combustion/mechanism.f90:17-108).  This tool never executes
them.  It parses the statements of

* ``PRODRATES``      (e.g. combustion/HYDROGENF.f90:4-277): third-body sums
  ``M(j)``, rate coefficients ``K(r)`` (Arrhenius, fitted reverse Arrhenius,
  Troe/Lindemann fall-off through ``GETLINDRATECOEFF``), rates of progress
  ``W(r)`` and production rates ``CDOT(s)``;
* ``COMPTHERMODATA`` (HYDROGENF.f90:514-673): pre-scaled NASA-7 polynomials;
* ``GETMOLARMASS`` / ``GETSPECIESNAMES`` (HYDROGENF.f90:366-405);
* ``GETMUCOEFF`` / ``GETKOVEREPS`` (HYDROGENF.f90:453-496), the transport fits;

and reproduces Fortran literal kinds exactly: the reference build sets no
default-real promotion (Makefile.in:19,23-24), so an unsuffixed real literal
such as ``-0.67`` or ``479604.2721`` is a REAL(4) value widened to double,
``2.64D+13`` and ``..._DP`` literals are exact doubles and integer literals are
exact.  Mixed-kind arithmetic follows Fortran promotion (int < real4 < real8);
real4-only sub-expressions are folded in single precision.

Output: a JSON intermediate form (``data/<mech>.json``, doubles stored as
``float.hex``) and a C header of static tables (``../csrc/mech_<mech>.h``)
shared by the HIP library and the CPU oracle.

Usage: ``python -m tiger_nga_amd.mech.mechc [--ref /root/reference]``
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys

import numpy as np

# name in this project -> (Fortran source, header) relative to the reference root
MECHANISMS = {
    "h2": ("combustion/HYDROGENF.f90", "combustion/HYDROGENF90.h"),
    "gri30": ("combustion/grimech30F.f90", "combustion/grimech30F90.h"),
    "dme39": ("combustion/DME39F.f90", "combustion/DME39F90.h"),
    "ethyred": ("combustion/EthyRedF.f90", "combustion/EthyRedF90.h"),
}

HERE = os.path.dirname(os.path.abspath(__file__))


# ----------------------------------------------------------------------------
# Source handling
# ----------------------------------------------------------------------------
def _strip_comment(line: str) -> str:
    out, q = [], None
    for ch in line:
        if q:
            out.append(ch)
            if ch == q:
                q = None
        elif ch in "'\"":
            q = ch
            out.append(ch)
        elif ch == "!":
            break
        else:
            out.append(ch)
    return "".join(out)


def logical_lines(text: str):
    """Free-form Fortran: drop comments, join ``&`` continuations.
    Yields (first physical line number, statement text)."""
    buf, start = "", None
    for no, raw in enumerate(text.splitlines(), 1):
        line = _strip_comment(raw).rstrip()
        if not line.strip():
            if buf:
                continue
            continue
        s = line.strip()
        if buf and s.startswith("&"):
            s = s[1:]
        cont = s.endswith("&")
        if cont:
            s = s[:-1]
        if not buf:
            start = no
        buf += " " + s
        if not cont:
            yield start, buf.strip()
            buf = ""
    if buf:
        yield start, buf.strip()


def header_params(text: str) -> dict:
    """``integer, parameter :: SN2 = 1, SH = 2, ...`` -> {'sn2': 1, ...}."""
    out = {}
    for _, st in logical_lines(text):
        m = re.match(r"integer\s*,\s*parameter\s*::(.*)$", st, re.I)
        if not m:
            continue
        for item in m.group(1).split(","):
            if "=" in item:
                k, v = item.split("=")
                v = v.strip()
                if re.fullmatch(r"\d+", v):
                    out[k.strip().lower()] = int(v)
    return out


def subroutine_body(lines, name: str, kind: str = "subroutine"):
    """Statements between ``subroutine NAME`` and the matching ``END``."""
    body, inside = [], False
    for no, st in lines:
        low = st.lower()
        if not inside:
            if re.match(r"(%s|double precision function)\s+%s\b" % (kind, name.lower()), low):
                inside = True
            continue
        if re.fullmatch(r"end(\s+(subroutine|function)(\s+\w+)?)?", low):
            return body
        body.append((no, st))
    raise ValueError("routine %s not found" % name)


# ----------------------------------------------------------------------------
# Expression parser with Fortran literal kinds
# ----------------------------------------------------------------------------
_TOK = re.compile(
    r"\s*(?:(?P<num>(?:\d+\.\d*|\.\d+|\d+)(?:[eEdD][+-]?\d+)?(?:_dp|_DP|_wp|_WP)?)"
    r"|(?P<name>[A-Za-z_][A-Za-z_0-9]*)|(?P<op>\*\*|[-+*/(),]))")


class Num:
    """Numeric literal or folded constant: ``kind`` in {'i', 'f4', 'f8'}."""
    __slots__ = ("v", "kind")

    def __init__(self, v, kind):
        self.v, self.kind = v, kind

    def __repr__(self):
        return "Num(%r,%s)" % (self.v, self.kind)


class Var:
    __slots__ = ("name",)

    def __init__(self, name):
        self.name = name.lower()

    def __repr__(self):
        return "Var(%s)" % self.name


class Ref:
    """Array element or function call ``name(args)``."""
    __slots__ = ("name", "args")

    def __init__(self, name, args):
        self.name, self.args = name.lower(), args

    def __repr__(self):
        return "Ref(%s,%r)" % (self.name, self.args)


class Bin:
    __slots__ = ("op", "a", "b")

    def __init__(self, op, a, b):
        self.op, self.a, self.b = op, a, b

    def __repr__(self):
        return "Bin(%s,%r,%r)" % (self.op, self.a, self.b)


class Neg:
    __slots__ = ("a",)

    def __init__(self, a):
        self.a = a

    def __repr__(self):
        return "Neg(%r)" % (self.a,)


def _literal(tok: str) -> Num:
    t = tok.lower()
    if t.endswith("_dp") or t.endswith("_wp"):
        return Num(float(t[:-3].replace("d", "e")), "f8")
    if "d" in t:
        return Num(float(t.replace("d", "e")), "f8")
    if "." in t or "e" in t:
        # default REAL: the decimal string rounded to binary32
        return Num(float(np.float32(t)), "f4")
    return Num(int(t), "i")


def tokenize(s: str):
    pos, out = 0, []
    while pos < len(s):
        if s[pos:].strip() == "":
            break
        m = _TOK.match(s, pos)
        if not m or m.end() == pos:
            raise SyntaxError("cannot tokenize %r at %d" % (s, pos))
        pos = m.end()
        if m.group("num"):
            out.append(("num", m.group("num")))
        elif m.group("name"):
            out.append(("name", m.group("name")))
        else:
            out.append(("op", m.group("op")))
    return out


class Parser:
    """Fortran precedence: ** > (* /) > unary +/- (applies to the first
    add-operand) > binary (+ -)."""

    def __init__(self, toks):
        self.t, self.i = toks, 0

    def peek(self):
        return self.t[self.i] if self.i < len(self.t) else (None, None)

    def take(self, val=None):
        tok = self.peek()
        if val is not None and tok[1] != val:
            raise SyntaxError("expected %s got %r" % (val, tok))
        self.i += 1
        return tok

    def expr(self):
        k, v = self.peek()
        if v in ("+", "-"):
            self.take()
            node = self.term()
            if v == "-":
                node = Neg(node)
        else:
            node = self.term()
        while self.peek()[1] in ("+", "-"):
            op = self.take()[1]
            node = Bin(op, node, self.term())
        return node

    def term(self):
        node = self.power()
        while self.peek()[1] in ("*", "/"):
            op = self.take()[1]
            node = Bin(op, node, self.power())
        return node

    def power(self):
        base = self.primary()
        if self.peek()[1] == "**":
            self.take()
            # right-associative; exponent may carry a sign
            k, v = self.peek()
            if v in ("+", "-"):
                self.take()
                e = self.power()
                e = Neg(e) if v == "-" else e
            else:
                e = self.power()
            return Bin("**", base, e)
        return base

    def primary(self):
        k, v = self.take()
        if k == "num":
            return _literal(v)
        if k == "name":
            if self.peek()[1] == "(":
                self.take("(")
                args = [self.expr()]
                while self.peek()[1] == ",":
                    self.take()
                    args.append(self.expr())
                self.take(")")
                return Ref(v, args)
            return Var(v)
        if v == "(":
            e = self.expr()
            self.take(")")
            return e
        if v in ("+", "-"):
            # a signed operand after an operator, e.g. TEMP / -10210
            # (EthyRedF.f90 Troe centres): the common compiler extension
            # binds the sign to the operand that follows
            e = self.power()
            return Neg(e) if v == "-" else e
        raise SyntaxError("unexpected %r" % v)


def parse_expr(s: str):
    p = Parser(tokenize(s))
    e = p.expr()
    if p.i != len(p.t):
        raise SyntaxError("trailing tokens in %r" % s)
    return e


def parse_assign(st: str):
    lhs, rhs = st.split("=", 1)
    return parse_expr(lhs), parse_expr(rhs)


# ----------------------------------------------------------------------------
# Constant folding with Fortran kind promotion
# ----------------------------------------------------------------------------
_RANK = {"i": 0, "f4": 1, "f8": 2}


def _as(v, kind):
    if kind == "i":
        return int(v)
    if kind == "f4":
        return np.float32(v)
    return float(v)


def fold(node):
    """Fold numeric sub-trees; returns Num or a rebuilt tree."""
    if isinstance(node, Num):
        return node
    if isinstance(node, Neg):
        a = fold(node.a)
        if isinstance(a, Num):
            return Num(-a.v, a.kind)
        return Neg(a)
    if isinstance(node, Bin):
        a, b = fold(node.a), fold(node.b)
        if isinstance(a, Num) and isinstance(b, Num):
            kind = max(a.kind, b.kind, key=_RANK.get)
            x, y = _as(a.v, kind), _as(b.v, kind)
            if node.op == "+":
                r = x + y
            elif node.op == "-":
                r = x - y
            elif node.op == "*":
                r = x * y
            elif node.op == "/":
                if kind == "i":
                    r = int(x / y)  # Fortran integer division truncates
                else:
                    r = x / y
            else:
                r = x ** y
            return Num(float(r) if kind != "i" else int(r), kind)
        return Bin(node.op, a, b)
    if isinstance(node, Ref):
        return Ref(node.name, [fold(x) for x in node.args])
    return node


def dval(n: Num) -> float:
    """Value of a literal as the double it becomes in mixed arithmetic."""
    if not isinstance(n, Num):
        raise TypeError("expected a constant, got %r" % (n,))
    return float(n.v)


# ----------------------------------------------------------------------------
# Pattern extraction
# ----------------------------------------------------------------------------
class MechError(ValueError):
    pass


def sum_terms(node, sign=1.0):
    """Flatten a +/- chain into [(sign, term)]."""
    if isinstance(node, Bin) and node.op in ("+", "-"):
        left = sum_terms(node.a, sign)
        right = sum_terms(node.b, sign if node.op == "+" else -sign)
        return left + right
    if isinstance(node, Neg):
        return sum_terms(node.a, -sign)
    return [(sign, node)]


def product_factors(node):
    if isinstance(node, Bin) and node.op == "*":
        return product_factors(node.a) + product_factors(node.b)
    return [node]


def exp_argument(node):
    """``b*LT + e/RT`` as (b, e) with the exact doubles of the source."""
    b = e = 0.0
    for sgn, t in sum_terms(node):
        t = fold(t)
        if isinstance(t, Bin) and t.op == "*" and isinstance(t.b, Var) and t.b.name == "lt":
            b += sgn * dval(t.a)
        elif isinstance(t, Bin) and t.op == "*" and isinstance(t.a, Var) and t.a.name == "lt":
            b += sgn * dval(t.b)
        elif isinstance(t, Var) and t.name == "lt":
            b += sgn
        elif isinstance(t, Bin) and t.op == "/" and isinstance(t.b, Var) and t.b.name == "rt":
            e += sgn * dval(t.a)
        else:
            raise MechError("unsupported exp() term %r" % (t,))
    return b, e


def arrhenius(node):
    """``A [* exp(b*LT + e/RT)]`` -> {'A','b','e'} (k = A*exp(b*lnT + e/RT))."""
    node = fold(node)
    if isinstance(node, Num):
        return {"A": dval(node), "b": 0.0, "e": 0.0}
    f = product_factors(node)
    if len(f) == 2 and isinstance(f[0], Num) and isinstance(f[1], Ref) and f[1].name == "exp":
        b, e = exp_argument(f[1].args[0])
        return {"A": dval(f[0]), "b": b, "e": e}
    raise MechError("unsupported rate expression %r" % (node,))


def troe_fc(node):
    """``c*EXP(-TEMP/d) + ... + c*EXP(-d/TEMP)`` -> [(c, kind, d)], kind 0:
    exp(-T/d), kind 1: exp(-d/T)."""
    out = []
    for sgn, t in sum_terms(node):
        t = fold(t)
        f = product_factors(t)
        if len(f) != 2 or not isinstance(f[0], Num) or not (isinstance(f[1], Ref) and f[1].name == "exp"):
            raise MechError("unsupported Fc term %r" % (t,))
        c = sgn * dval(f[0])
        arg = fold(f[1].args[0])
        if not isinstance(arg, Neg) or not (isinstance(arg.a, Bin) and arg.a.op == "/"):
            raise MechError("unsupported Fc exponent %r" % (arg,))
        num, den = arg.a.a, arg.a.b
        if isinstance(num, Var) and num.name == "temp" and isinstance(den, Num):
            out.append((c, 0, dval(den)))
        elif isinstance(den, Var) and den.name == "temp" and isinstance(num, Num):
            out.append((c, 1, dval(num)))
        else:
            raise MechError("unsupported Fc exponent %r" % (arg,))
    return out


def _idx(node, params):
    if isinstance(node, Var):
        return params[node.name] - 1
    if isinstance(node, Num):
        return int(node.v) - 1
    raise MechError("bad index %r" % (node,))


def parse_prodrates(body, params):
    third = {}      # j -> {species: coef}
    slots = {}      # slot -> rate description
    wfac = {}       # slot -> {'c': [species...], 'm': j or -1}
    cdot = {}       # species -> {slot: coef}
    k0 = kinf = fc = None
    src = {}        # slot -> first physical line of K assignment
    for no, st in body:
        low = st.lower()
        if low.startswith(("implicit", "include", "real", "integer")) or "=" not in st:
            continue
        lhs, rhs = parse_assign(st)
        if isinstance(lhs, Var) and lhs.name in ("lt", "rt"):
            continue
        if isinstance(lhs, Var) and lhs.name == "k0troe":
            k0 = arrhenius(rhs)
            continue
        if isinstance(lhs, Var) and lhs.name == "kinftroe":
            kinf = arrhenius(rhs)
            continue
        if isinstance(lhs, Var) and lhs.name == "fctroe":
            fc = troe_fc(rhs)
            continue
        if not isinstance(lhs, Ref):
            raise MechError("line %d: unsupported statement %s" % (no, st))
        i = _idx(lhs.args[0], params)
        if lhs.name == "m":
            acc = third.setdefault(i, {}) if True else None
            for sgn, t in sum_terms(rhs):
                t = fold(t)
                if isinstance(t, Ref) and t.name == "m" and _idx(t.args[0], params) == i:
                    continue  # M(j) = M(j) + ... accumulation
                f = product_factors(t)
                coef = 1.0
                if len(f) == 2 and isinstance(f[0], Num):
                    coef, f = dval(f[0]), f[1:]
                if len(f) != 1 or not (isinstance(f[0], Ref) and f[0].name == "c"):
                    raise MechError("line %d: bad third-body term %r" % (no, t))
                s = _idx(f[0].args[0], params)
                acc[s] = acc.get(s, 0.0) + sgn * coef
        elif lhs.name == "k":
            src[i] = no
            r = fold(rhs)
            if isinstance(r, Ref) and r.name == "getlindratecoeff":
                a = r.args
                if not (isinstance(a[2], Var) and a[2].name == "k0troe" and isinstance(a[3], Var)
                        and a[3].name == "kinftroe" and isinstance(a[4], Var) and a[4].name == "fctroe"):
                    raise MechError("line %d: unexpected fall-off call" % no)
                if isinstance(a[5], Ref) and a[5].name == "m":
                    mj = _idx(a[5].args[0], params)
                elif isinstance(a[5], Var) and a[5].name == "concdefault":
                    mj = -1
                else:
                    raise MechError("line %d: bad fall-off concentration" % no)
                if k0 is None or kinf is None or fc is None:
                    raise MechError("line %d: fall-off before its parameters" % no)
                slots[i] = {"type": "falloff", "k0": dict(k0), "kinf": dict(kinf),
                            "fc": [list(x) for x in fc], "m": mj}
                # GETLINDRATECOEFF overwrites its K0 argument (HYDROGENF.f90:302);
                # every shipped mechanism reassigns K0TROE before the next use.
                k0 = None
            else:
                slots[i] = dict(type="arrhenius", **arrhenius(r))
        elif lhs.name == "w":
            f = product_factors(fold(rhs))
            if not (isinstance(f[0], Ref) and f[0].name == "k" and _idx(f[0].args[0], params) == i):
                raise MechError("line %d: W must start with its own K" % no)
            cs, mj = [], -1
            for x in f[1:]:
                if isinstance(x, Ref) and x.name == "c":
                    cs.append(_idx(x.args[0], params))
                elif isinstance(x, Ref) and x.name == "m" and mj < 0:
                    mj = _idx(x.args[0], params)
                else:
                    raise MechError("line %d: bad W factor %r" % (no, x))
            wfac[i] = {"c": cs, "m": mj}
        elif lhs.name == "cdot":
            acc = cdot.setdefault(i, {})
            r = fold(rhs)
            if isinstance(r, Num) and r.v == 0:
                continue
            for sgn, t in sum_terms(r):
                t = fold(t)
                if isinstance(t, Ref) and t.name == "cdot" and _idx(t.args[0], params) == i:
                    continue  # CDOT(s) = CDOT(s) + ... accumulation
                f = product_factors(t)
                coef = 1.0
                if len(f) == 2 and isinstance(f[0], Num):
                    coef, f = dval(f[0]), f[1:]
                if len(f) != 1 or not (isinstance(f[0], Ref) and f[0].name == "w"):
                    raise MechError("line %d: bad CDOT term %r" % (no, t))
                sl = _idx(f[0].args[0], params)
                acc[sl] = acc.get(sl, 0.0) + sgn * coef
        else:
            raise MechError("line %d: unsupported statement %s" % (no, st))
    return third, slots, wfac, cdot, src


def horner(node):
    """``c1 + T*(c2 + T*(...))`` -> [c1, c2, ...]."""
    node = fold(node)
    if isinstance(node, Num):
        return [dval(node)]
    if isinstance(node, Bin) and node.op in ("+", "-") and isinstance(node.a, Num):
        rest = node.b
        if not (isinstance(rest, Bin) and rest.op == "*" and isinstance(rest.a, Var) and rest.a.name == "t"):
            raise MechError("bad Horner form %r" % (node,))
        inner = horner(rest.b)
        if node.op == "-":
            inner = [-x for x in inner]
        return [dval(node.a)] + inner
    raise MechError("bad Horner form %r" % (node,))


def parse_thermo(body, params, nsp):
    """COMPTHERMODATA: ``IF (T.GT.tmid) THEN <high> ELSE <low> END IF``."""
    tmid, block = None, None
    out = {"hi": {}, "lo": {}}
    for no, st in body:
        low = st.lower().replace(" ", "")
        m = re.match(r"if\(t\.gt\.([0-9.e+\-_dp]+)\)then", low)
        if m:
            tmid, block = dval(_literal(m.group(1))), "hi"
            continue
        if low == "else":
            block = "lo"
            continue
        if low in ("endif",) or low.startswith(("implicit", "include", "real")):
            continue
        if block is None or "=" not in st:
            continue
        lhs, rhs = parse_assign(st)
        s = _idx(lhs.args[0], params)
        f = fold(rhs)
        if not (isinstance(f, Bin) and f.op == "*" and isinstance(f.a, Num)):
            raise MechError("line %d: bad thermo form" % no)
        a, inner = dval(f.a), f.b
        ent = out[block].setdefault(s, {})
        if lhs.name == "h":
            # a * ( T * (poly) +/- c6 )
            if not (isinstance(inner, Bin) and inner.op in ("+", "-") and isinstance(inner.b, Num)):
                raise MechError("line %d: bad enthalpy form" % no)
            tp = inner.a
            if not (isinstance(tp, Bin) and tp.op == "*" and isinstance(tp.a, Var) and tp.a.name == "t"):
                raise MechError("line %d: bad enthalpy form" % no)
            c = horner(tp.b)
            c6 = dval(inner.b) * (1.0 if inner.op == "+" else -1.0)
            ent["h"] = [a] + c + [c6]
        elif lhs.name == "cp":
            ent["cp"] = [a] + horner(inner)
        else:
            raise MechError("line %d: unexpected thermo target" % no)
    for blk in ("hi", "lo"):
        if sorted(out[blk]) != list(range(nsp)):
            raise MechError("thermo block %s incomplete" % blk)
        for s in range(nsp):
            if len(out[blk][s]["h"]) != 7 or len(out[blk][s]["cp"]) != 6:
                raise MechError("species %d: expected NASA-7 degree-5 forms" % s)
    return tmid, out


def compile_mechanism(name: str, ref_root: str) -> dict:
    f90, h = MECHANISMS[name]
    text = open(os.path.join(ref_root, f90)).read()
    params = header_params(open(os.path.join(ref_root, h)).read())
    lines = list(logical_lines(text))
    nsp = params["send"] - 1
    nslot = params["rend"] - 1 if "rend" in params else None
    nm = params["mend"] - 1

    names = [None] * nsp
    for no, st in subroutine_body(lines, "GETSPECIESNAMES"):
        m = re.match(r"NAMES\((\w+)\)\s*=\s*'([^']*)'", st, re.I)
        if m:
            names[params[m.group(1).lower()] - 1] = m.group(2).strip()
    wmol = [None] * nsp
    for no, st in subroutine_body(lines, "GETMOLARMASS"):
        if re.match(r"MM\(", st, re.I):
            lhs, rhs = parse_assign(st)
            wmol[_idx(lhs.args[0], params)] = dval(fold(rhs))
    # transport fits (GETMUCOEFF / GETKOVEREPS, e.g. HYDROGENF.f90:453-496):
    # unsuffixed literals, so binary32 values widened to double
    mucoeff, koveps = [None] * nsp, [None] * nsp
    for sub, arr, var in (("GETMUCOEFF", mucoeff, "MUCOEFF"), ("GETKOVEREPS", koveps, "KOVEREPS")):
        for no, st in subroutine_body(lines, sub):
            if re.match(var + r"\(", st, re.I):
                lhs, rhs = parse_assign(st)
                arr[_idx(lhs.args[0], params)] = dval(fold(rhs))
    if None in mucoeff or None in koveps:
        raise MechError("incomplete transport data")
    nreac_decl = None
    for no, st in subroutine_body(lines, "GETNREACTIONS"):
        m = re.match(r"NREACTIONS\s*=\s*(\d+)", st, re.I)
        if m:
            nreac_decl = int(m.group(1))
    if None in names or None in wmol:
        raise MechError("incomplete species data")

    third, slots, wfac, cdot, src = parse_prodrates(subroutine_body(lines, "PRODRATES"), params)
    tmid, thermo = parse_thermo(subroutine_body(lines, "COMPTHERMODATA"), params, nsp)

    nslot = max(slots) + 1
    if sorted(slots) != list(range(nslot)) or sorted(wfac) != list(range(nslot)):
        raise MechError("rate slots are not contiguous")
    if nreac_decl is not None and nreac_decl != nslot:
        raise MechError("GETNREACTIONS=%d but %d slots" % (nreac_decl, nslot))

    # slot names from the header: R<n>F / R<n>B / R<n> (and RG<n>...)
    slotname = {}
    for k, v in params.items():
        # every R-named parameter below REND is a rate slot: R12, R2F / R2B,
        # and free-form names such as RG04, RLPP009F, RIC67 (EthyRedF90.h)
        if k.startswith("r") and k != "rend" and 1 <= v < params.get("rend", 1 << 30):
            slotname[v - 1] = k.upper()
    # pair forward/backward slots into reactions
    reactions = []
    s = 0
    while s < nslot:
        nm_ = slotname[s]
        if nm_.endswith("F"):
            base = nm_[:-1]
            if slotname.get(s + 1) != base + "B":
                raise MechError("forward slot %s without backward" % nm_)
            pair = (s, s + 1)
            s += 2
        else:
            base, pair = nm_, (s, -1)
            s += 1
        reactions.append((base, pair))

    rx_out = []
    for base, (sf, sb) in reactions:
        rf = slots[sf]
        rb = slots[sb] if sb >= 0 else None
        nu = {}
        for sp, row in cdot.items():
            cf = row.get(sf, 0.0)
            cb = row.get(sb, 0.0) if sb >= 0 else 0.0
            if sb >= 0 and cf != -cb:
                raise MechError("%s: species %s not antisymmetric in F/B" % (base, names[sp]))
            if cf != 0.0:
                nu[sp] = cf
        ent = {
            "name": base, "slot_f": sf, "slot_b": sb, "line": src[sf],
            "reac": wfac[sf]["c"], "prod": wfac[sb]["c"] if sb >= 0 else [],
            "m_w": wfac[sf]["m"],
            "nu": sorted([[sp, c] for sp, c in nu.items()]),
        }
        if sb >= 0 and wfac[sb]["m"] != wfac[sf]["m"]:
            raise MechError("%s: F/B third-body mismatch" % base)
        if len(ent["reac"]) > 3 or len(ent["prod"]) > 3:
            raise MechError("%s: more than 3 concentration factors" % base)
        if rf["type"] == "falloff":
            if rb is not None and (rb["type"] != "falloff" or rb["fc"] != rf["fc"] or rb["m"] != rf["m"]):
                raise MechError("%s: backward fall-off does not share Fc/M" % base)
            ent.update(type="falloff", m_fo=rf["m"], fc=rf["fc"],
                       f_k0=rf["k0"], f_kinf=rf["kinf"],
                       b_k0=rb["k0"] if rb else None, b_kinf=rb["kinf"] if rb else None)
            if ent["m_w"] >= 0:
                raise MechError("%s: fall-off rate with extra third-body factor" % base)
        else:
            if rb is not None and rb["type"] != "arrhenius":
                raise MechError("%s: mixed F/B rate kinds" % base)
            ent.update(type="third_body" if ent["m_w"] >= 0 else "elementary",
                       f={k: rf[k] for k in "Abe"}, b={k: rb[k] for k in "Abe"} if rb else None)
        rx_out.append(ent)

    eff = [[0.0] * nsp for _ in range(nm)]
    for j, row in third.items():
        for sp, c in row.items():
            eff[j][sp] = c

    return {
        "name": name, "source": f90, "header": h,
        "species": names, "W": wmol, "mucoeff": mucoeff, "koveps": koveps, "nslot": nslot, "n_third": nm,
        "third_body": eff, "tmid": tmid,
        "thermo_hi": [thermo["hi"][s] for s in range(nsp)],
        "thermo_lo": [thermo["lo"][s] for s in range(nsp)],
        "reactions": rx_out,
        "cdot_slots": {str(sp): sorted([[sl, c] for sl, c in row.items()]) for sp, row in cdot.items()},
    }


# ----------------------------------------------------------------------------
# Emitters
# ----------------------------------------------------------------------------
def _hex(x):
    return float(x).hex()


def to_json(mech: dict) -> str:
    def conv(o):
        if isinstance(o, float):
            return {"hex": _hex(o), "v": o}
        if isinstance(o, dict):
            return {k: conv(v) for k, v in o.items()}
        if isinstance(o, (list, tuple)):
            return [conv(v) for v in o]
        return o
    return json.dumps(conv(mech), indent=1)


def from_json(text: str) -> dict:
    def conv(o):
        if isinstance(o, dict):
            if set(o) == {"hex", "v"}:
                return float.fromhex(o["hex"])
            return {k: conv(v) for k, v in o.items()}
        if isinstance(o, list):
            return [conv(v) for v in o]
        return o
    return conv(json.loads(text))


def _c(x):
    """Exact C double literal."""
    x = float(x)
    if x == 0.0:
        return "0.0"
    return _hex(x)


def emit_c_header(mech: dict) -> str:
    """Static tables in the layout of ``tgr_mech_tables`` (mech_tables.h)."""
    n, P = mech["name"], mech["name"].upper()
    nsp, rx = len(mech["species"]), mech["reactions"]
    L = []
    w = L.append
    w("/* Generated by tiger_nga_amd/mech/mechc.py from the reference mechanism")
    w(" * %s (+ %s).  Do not edit.  Exact doubles are hex literals;" % (mech["source"], mech["header"]))
    w(" * unsuffixed Fortran reals carry their binary32 value. */")
    w("#pragma once")
    w('#include "mech_tables.h"')
    w("")
    w("static const char *const %s_species[%d] = {%s};" % (n, nsp, ", ".join('"%s"' % s for s in mech["species"])))
    w("static const double %s_W[%d] = {%s};" % (n, nsp, ", ".join(_c(x) for x in mech["W"])))
    w("static const double %s_mucoeff[%d] = {%s};" % (n, nsp, ", ".join(_c(x) for x in mech["mucoeff"])))
    w("static const double %s_koveps[%d] = {%s};" % (n, nsp, ", ".join(_c(x) for x in mech["koveps"])))
    # element composition (C, H, O, N, AR) from the species names
    from .meta import ELEMENTS, composition
    erows = []
    for el in ELEMENTS:
        erows.append("{" + ", ".join("%d" % composition(sp).get(el, 0) for sp in mech["species"]) + "}")
    w("static const double %s_elem[%d][%d] = {\n  %s};" % (n, len(ELEMENTS), nsp, ",\n  ".join(erows)))
    # thermo: [range][species][13]: a_h, c1..c5, c6, a_cp, d1..d5
    rows = []
    for blk in ("thermo_hi", "thermo_lo"):
        for s in range(nsp):
            t = mech[blk][s]
            rows.append("{" + ", ".join(_c(x) for x in t["h"] + t["cp"]) + "}")
    w("static const double %s_thermo[%d][13] = {\n  %s};" % (n, 2 * nsp, ",\n  ".join(rows)))
    nm = mech["n_third"]
    eff = ",\n  ".join("{" + ", ".join(_c(x) for x in row) + "}" for row in mech["third_body"]) if nm else ""
    w("static const double %s_eff[%d][%d] = {\n  %s};" % (n, max(nm, 1), nsp, eff or "{0}"))
    # reactions
    ents, nus = [], []
    for r in rx:
        kind = {"elementary": 0, "third_body": 1, "falloff": 2}[r["type"]]
        rev = 1 if r["slot_b"] >= 0 else 0
        reac = (r["reac"] + [-1, -1, -1])[:3]
        prod = (r["prod"] + [-1, -1, -1])[:3]
        if kind == 2:
            fk, fk0 = r["f_kinf"], r["f_k0"]
            bk = r["b_kinf"] or {"A": 0.0, "b": 0.0, "e": 0.0}
            bk0 = r["b_k0"] or {"A": 0.0, "b": 0.0, "e": 0.0}
            mi = r["m_fo"]
            fc = (r["fc"] + [[0.0, 0, 0.0]] * 3)[:3]
            nfc = len(r["fc"])
        else:
            fk, bk = r["f"], r["b"] or {"A": 0.0, "b": 0.0, "e": 0.0}
            fk0 = bk0 = {"A": 0.0, "b": 0.0, "e": 0.0}
            mi = r["m_w"]
            fc, nfc = [[0.0, 0, 0.0]] * 3, 0
        lind = 1 if (kind == 2 and nfc == 1 and fc[0][0] == 1.0 and fc[0][1] == 1 and fc[0][2] == 0.0) else 0
        nu0 = len(nus)
        for sp, c in r["nu"]:
            nus.append("{%d, %s}" % (sp, _c(c)))
        ents.append(
            "{%d, %d, %d, %d, %d, %d, {%s}, {%s}, %d, %d, %d, %d, %d,\n"
            "   {%s, %s, %s}, {%s, %s, %s}, {%s, %s, %s}, {%s, %s, %s},\n"
            "   {%s}, {%s}, {%s}}"
            % (kind, rev, mi, len(r["reac"]), len(r["prod"]), nfc,
               ", ".join(map(str, reac)), ", ".join(map(str, prod)),
               nu0, len(r["nu"]), r["slot_f"], r["slot_b"], lind,
               _c(fk["A"]), _c(fk["b"]), _c(fk["e"]), _c(bk["A"]), _c(bk["b"]), _c(bk["e"]),
               _c(fk0["A"]), _c(fk0["b"]), _c(fk0["e"]), _c(bk0["A"]), _c(bk0["b"]), _c(bk0["e"]),
               ", ".join(_c(x[0]) for x in fc), ", ".join(str(int(x[1])) for x in fc),
               ", ".join(_c(x[2]) for x in fc)))
    w("static const tgr_nu %s_nu[%d] = {\n  %s};" % (n, len(nus), ",\n  ".join(nus)))
    w("static const tgr_rxn %s_rxn[%d] = {\n  %s};" % (n, len(ents), ",\n  ".join(ents)))
    w("")
    w("static const tgr_mech_tables %s_tables = {" % n)
    w('  "%s", %d, %d, %d, %d, %d, %s,' % (n, nsp, len(rx), mech["nslot"], nm, len(nus), _c(mech["tmid"])))
    w("  %s_species, %s_W, &%s_thermo[0][0], &%s_eff[0][0], %s_rxn, %s_nu, %s_mucoeff, %s_koveps, &%s_elem[0][0]};"
      % (n, n, n, n, n, n, n, n, n))
    return "\n".join(L) + "\n"


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--ref", default="/root/reference")
    ap.add_argument("--mech", nargs="*", default=list(MECHANISMS))
    args = ap.parse_args(argv)
    os.makedirs(os.path.join(HERE, "data"), exist_ok=True)
    for name in args.mech:
        mech = compile_mechanism(name, args.ref)
        with open(os.path.join(HERE, "data", name + ".json"), "w") as f:
            f.write(to_json(mech))
        with open(os.path.join(HERE, "..", "csrc", "mech_%s.h" % name), "w") as f:
            f.write(emit_c_header(mech))
        nfo = sum(r["type"] == "falloff" for r in mech["reactions"])
        ntb = sum(r["type"] == "third_body" for r in mech["reactions"])
        print("%s: %d species, %d reactions (%d slots), %d third-body sums, %d third-body, %d fall-off"
              % (name, len(mech["species"]), len(mech["reactions"]), mech["nslot"], mech["n_third"], ntb, nfo))


if __name__ == "__main__":
    sys.exit(main())
