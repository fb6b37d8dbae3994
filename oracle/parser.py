"""Python restatement of gopf's equation parser (TEST INFRASTRUCTURE ONLY).

This module is part of the oracle: only tests, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` import it.  It restates, statement by
statement, how gopf turns equation strings into right-hand-side terms:

  pf/model.go:157-195      AddEquation / UpdateDerivedFields
  pf/model.go:197-234      AllFieldNames / IsFieldName / IsBrickName
  pf/rhsBuilder.go:26-106  Build / fieldNameFromLeibniz / isBilinear
  pf/rhsBuilder.go:108-190 ValidName / ConcreteTerm
  pf/rhsBuilder.go:199-287 constructFunc / known prefixes
  pf/util.go:17-104        GetNonLinearFieldExpressions / DerivedFieldCalcFromDesc /
                           GetPower / GetFieldName
  pf/util.go:138-213       SplitOnMany / removeEmpty
  pf/util.go:294-299       SortFactors

Go's ``regexp`` package is leftmost-first like Python's ``re`` for the patterns
used here; the one behavioural difference that matters -- Go's FindAll skips an
empty match that abuts the previous match -- is restated in ``go_find_all``.
Go's ``cmplx.Pow`` and complex multiplication are restated so scalar
coefficients carry the same rounding as the reference closures.
"""
from __future__ import annotations

import numpy as np

import cmath
import math
import re
from dataclasses import dataclass, field as dc_field
from typing import Dict, List, Optional, Tuple


class ParseError(Exception):
    """Raised where the reference panics (log.Fatalf / panic / runtime error)."""


# ---------------------------------------------------------------------------
# Go library restatements
# ---------------------------------------------------------------------------

def go_find_all(pattern: str, s: str) -> List[str]:
    """regexp.FindAllString(s, -1) (Go regexp.go allMatches)."""
    try:
        rx = re.compile(pattern)
    except re.error as exc:  # regexp.MustCompile panics
        raise ParseError(f"invalid regexp {pattern!r}: {exc}") from exc
    out: List[str] = []
    pos, prev_end, end = 0, -1, len(s)
    while pos <= end:
        m = rx.search(s, pos)
        if m is None:
            break
        accept = True
        if m.end() == pos:
            if m.start() == prev_end:
                accept = False
            pos = pos + 1 if pos < end else end + 1
        else:
            pos = m.end()
        prev_end = m.end()
        if accept:
            out.append(m.group(0))
    return out


def go_find_string(pattern: str, s: str) -> str:
    try:
        m = re.compile(pattern).search(s)
    except re.error as exc:
        raise ParseError(f"invalid regexp {pattern!r}: {exc}") from exc
    return m.group(0) if m else ""


def go_find_submatch(pattern: str, s: str) -> List[str]:
    m = re.compile(pattern).search(s)
    if not m:
        return []
    return [m.group(0)] + [g if g is not None else "" for g in m.groups()]


def go_parse_float(s: str) -> float:
    """strconv.ParseFloat for the digit strings GetPower extracts."""
    try:
        return float(s)
    except ValueError as exc:
        raise ParseError(f"strconv.ParseFloat: parsing {s!r}") from exc


def go_cmul(x: complex, y: complex) -> complex:
    a, b, c, d = x.real, x.imag, y.real, y.imag
    return complex(a * c - b * d, a * d + b * c)


def go_cpow(x: complex, p: float) -> complex:
    """math/cmplx.Pow(x, complex(p, 0))."""
    if x == 0:
        if p == 0:
            return complex(1.0, 0.0)
        if p < 0:
            return complex(math.inf, 0.0)
        return complex(0.0, 0.0)
    modulus = math.hypot(x.real, x.imag)
    if modulus == 0:
        return complex(0.0, 0.0)
    r = math.pow(modulus, p)
    theta = p * math.atan2(x.imag, x.real)
    return complex(r * math.cos(theta), r * math.sin(theta))


# ---------------------------------------------------------------------------
# pf/util.go restatements
# ---------------------------------------------------------------------------

@dataclass
class SubStringDelimiter:
    SubString: str
    PreceedingDelimiter: str


def get_first_delimiter(value: str, delimiters: List[str]) -> str:
    if value == "":
        raise ParseError("index out of range [0] with length 0")
    for d in delimiters:
        if value[0] == d:
            return d
    return ""


def split_on_many(value: str, delimiters: List[str]) -> List[SubStringDelimiter]:
    """SplitOnMany (pf/util.go:166-198), breadth-first queue order."""
    substrings: List[SubStringDelimiter] = []
    queue = [SubStringDelimiter(value, get_first_delimiter(value, delimiters))]
    all_delims = "".join(delimiters)
    while queue:
        cur = queue.pop(0)
        if not any(ch in cur.SubString for ch in all_delims):
            substrings.append(cur)
            continue
        delim = delimiters[0]
        for d in delimiters:
            if d in cur.SubString:
                delim = d
                break
        splits = [s for s in cur.SubString.split(delim) if s != ""]
        if not splits:
            raise ParseError("index out of range [0] with length 0")
        queue.append(SubStringDelimiter(splits[0], cur.PreceedingDelimiter))
        for s in splits[1:]:
            queue.append(SubStringDelimiter(s, delim))
    return substrings


def sort_factors(expr: str) -> str:
    return "*".join(sorted(expr.split("*")))


def get_power(pattern: str) -> float:
    m = go_find_submatch(r"\^(-?\d+\.?\d*)", pattern)
    if len(m) <= 1:
        return 1.0
    return go_parse_float(m[1])


def get_field_name(term: str, field_names: List[str]) -> str:
    fld = ""
    for f in field_names:
        if f in term:
            without = term.replace(f, "")
            ok = True
            for f1 in field_names:
                if f1 in without:
                    ok = False
                    break
            if ok and len(f) > len(fld):
                fld = f
    return fld


def get_non_linear_field_expressions(pattern: str, fld: str, field_names: List[str]) -> str:
    expr = ""
    for name in field_names:
        if name == fld and is_bilinear(pattern, fld, field_names):
            continue
        res = go_find_string(name + r"[^\*]*", pattern)
        if res != "":
            expr += res + "*"
    if len(expr) > 1:
        return expr[:-1]
    return expr


# ---------------------------------------------------------------------------
# pf/rhsBuilder.go restatements
# ---------------------------------------------------------------------------

KNOWN_PREFIXES = ["-", "LAP^4", "LAP^2", "LAP", "*"]


def field_name_from_leibniz(leibniz: str) -> str:
    if len(leibniz) <= 3:
        raise ParseError("rhsbuilder: Length of the Leibniz formatted string has to be at least 3")
    if leibniz[0:1] != "d" or leibniz[-3:] != "/dt":
        raise ParseError("rhsbuilder: Passed string is not a leibniz formatted string")
    return leibniz[1:-3]


def is_bilinear(term: str, fld: str, field_names: List[str]) -> bool:
    if len(go_find_all(fld, term)) != 1:
        return False
    for f in field_names:
        if f == fld:
            continue
        if len(go_find_all(f, term)) > 0:
            return False
    res = go_find_string(fld + r"*[^/\*]*", term)
    m = go_find_submatch(r"\^(-?\d+\.?\d*)", res)
    if len(m) <= 1:
        return True
    try:
        power = float(m[1])
    except ValueError:
        return True
    return abs(power - 1.0) < 1e-10


def remove_known_prefixes(s: str) -> str:
    prefixes = list(KNOWN_PREFIXES)
    while prefixes:
        p = prefixes.pop(0)
        if s.startswith(p):
            s = s[len(p):]
            prefixes = list(KNOWN_PREFIXES)
    return s


def get_known_prefixes(s: str) -> List[str]:
    pref: List[str] = []
    prefixes = list(KNOWN_PREFIXES)
    while prefixes:
        p = prefixes.pop(0)
        if s.startswith(p):
            pref.append(p)
            prefixes = list(KNOWN_PREFIXES)
            s = s[len(p):]
    return pref


# ---------------------------------------------------------------------------
# Model (pf/model.go) restatement producing a flat term description
# ---------------------------------------------------------------------------

KIND_CONCRETE, KIND_SQGRAD, KIND_ELAST, KIND_VOLCONS, KIND_SVISC, KIND_HESS, KIND_CNOISE = 0, 1, 2, 3, 4, 5, 6
_USER_KINDS = (KIND_SQGRAD, KIND_ELAST, KIND_VOLCONS, KIND_SVISC, KIND_HESS, KIND_CNOISE)


@dataclass
class TermDesc:
    """One closure of RHS.Terms / RHS.Denum, as data."""
    kind: int
    coef: complex = 1 + 0j          # concrete: sign * prod Pow(scalar, p); user: +-1
    brick: int = -1                 # concrete: brick index (fields then derived) or -1
    has_lap: bool = False
    lap: int = 0                    # concrete: LaplacianN power; user: total LAP prefix power
    user: int = -1                  # user term table index
    sign: float = 1.0               # concrete: sign of the term
    factors: List[Tuple[str, float]] = dc_field(default_factory=list)  # concrete: (scalar, power)

    def coef_derivative(self, scalars: Dict[str, complex], name: str) -> float:
        """d coef / d scalar for real scalars: sign * p s^(p-1) * the other factors (no reference
        counterpart; the parameter gradient of the adjoint)."""
        d = 0.0
        for j, (nm, p) in enumerate(self.factors):
            if nm != name:
                continue
            v = self.sign * p * scalars[name].real ** (p - 1.0)
            for i, (other, q) in enumerate(self.factors):
                if i != j:
                    v *= scalars[other].real ** q
            d += v
        return d


@dataclass
class DerivedDesc:
    name: str
    kind: int                       # 0 product of cmplx.Pow factors, 1 real polynomial, 2 Gaussian noise
    terms: List[Tuple[float, List[Tuple[int, float]]]] = dc_field(default_factory=list)
    noise_std: float = 0.0
    noise_seed: int = 0


@dataclass
class UserTerm:
    kind: int                       # KIND_SQGRAD, KIND_ELAST, KIND_VOLCONS or KIND_SVISC
    params: dict


class Model:
    """Restatement of pf.Model's bookkeeping (model.go:116-260)."""

    def __init__(self) -> None:
        self.fields: List[Tuple[str, int]] = []       # (name, n)
        self.derived: List[DerivedDesc] = []
        self.scalars: Dict[str, complex] = {}
        self.implicit_terms: Dict[str, UserTerm] = {}
        self.explicit_terms: Dict[str, UserTerm] = {}
        self.equations: List[str] = []
        self.rhs: List[Tuple[List[TermDesc], List[TermDesc]]] = []
        self.user_tables: Dict[int, List[dict]] = {k: [] for k in _USER_KINDS}

    # -- names -------------------------------------------------------------
    def field_names(self) -> List[str]:
        return [f[0] for f in self.fields]

    def all_field_names(self) -> List[str]:
        return self.field_names() + [d.name for d in self.derived]

    def is_field_name(self, name: str) -> bool:
        return name in self.all_field_names()

    def is_brick_name(self, name: str) -> bool:
        return name in self.all_field_names() or name in self.scalars

    def brick_index(self, name: str) -> int:
        return self.all_field_names().index(name)

    def is_user_defined_term(self, desc: str) -> bool:
        return desc in self.implicit_terms or desc in self.explicit_terms

    # -- model building ------------------------------------------------------
    @staticmethod
    def _panic_on_prefix_in_name(name: str) -> None:
        if len(get_known_prefixes(name)) > 0:
            raise ParseError(f"The words {KNOWN_PREFIXES} are reserved. Do not include them in your variable names")

    def add_field(self, name: str, n: int) -> None:
        self._panic_on_prefix_in_name(name)
        self.fields.append((name, n))

    def add_scalar(self, name: str, value: complex) -> None:
        self._panic_on_prefix_in_name(name)
        self.scalars[name] = complex(value)

    def register_explicit_term(self, name: str, kind: int, **params) -> None:
        self._panic_on_prefix_in_name(name)
        self.explicit_terms[name] = UserTerm(kind, params)

    def register_implicit_term(self, name: str, kind: int, **params) -> None:
        self._panic_on_prefix_in_name(name)
        self.implicit_terms[name] = UserTerm(kind, params)

    def register_polynomial(self, name: str, terms: List[Tuple[float, Dict[str, float]]]) -> None:
        """RegisterFunction (model.go:400-412) with a polynomial of the fields' real parts."""
        self._panic_on_prefix_in_name(name)
        idx = {n: i for i, n in enumerate(self.field_names())}
        dterms = []
        for coef, powers in terms:
            dterms.append((float(coef), [(idx[f], float(p)) for f, p in powers.items()]))
        self.derived.append(DerivedDesc(name, 1, dterms))

    def register_white_noise(self, name: str, strength: float, seed: int) -> None:
        """RegisterFunction(name, WhiteNoise{strength}.Generate) (noise.go:14-22); the draws come
        from a counter-based generator keyed by seed."""
        self._panic_on_prefix_in_name(name)
        self.derived.append(DerivedDesc(name, 2, [], float(np.sqrt(2.0 * strength)), int(seed)))

    def register_conservative_noise(self, term: str, strength: float, dim: int, seed: int) -> None:
        """NewConservativeNoise + RegisterExplicitTerm with RequiredDerivedFields (noise.go:24-100)."""
        self._panic_on_prefix_in_name(term)
        prefix = seed & 0xFFFFFFFF
        names = []
        for c in range(dim):
            name = f"{prefix}_current_{c}"
            s = (seed + 0x9E3779B97F4A7C15 * (c + 1)) & 0xFFFFFFFFFFFFFFFF
            self.derived.append(DerivedDesc(name, 2, [], float(np.sqrt(2.0 * strength)), s))
            names.append(name)
        self.explicit_terms[term] = UserTerm(KIND_CNOISE, {"names": names})

    def add_equation(self, eq: str) -> None:
        eq = eq.replace(" ", "")
        self.equations.append(eq)
        self._update_derived_fields(eq)

    def _update_derived_fields(self, eq: str) -> None:
        parts = eq.split("=")
        if len(parts) < 2:
            raise ParseError("index out of range [1] with length 1")
        rhs = parts[1]
        splitted = [s.SubString for s in split_on_many(rhs, ["+", "-"])]
        fld = field_name_from_leibniz(parts[0])
        fnames = self.field_names()
        for s in splitted:
            if self.is_user_defined_term(s):
                continue
            new = sort_factors(get_non_linear_field_expressions(s, fld, fnames))
            if new != "" and not self.is_field_name(new):
                if not self.fields:
                    raise ParseError("index out of range [0] with length 0")
                self.derived.append(self._derived_from_desc(new))

    def _derived_from_desc(self, desc: str) -> DerivedDesc:
        """DerivedFieldCalcFromDesc (util.go:38-65)."""
        idx = {n: i for i, n in enumerate(self.field_names())}
        factors = []
        for part in go_find_all(r"[^\*]*", desc):
            name = go_find_string(r"^[^\^]*", part)
            if name not in idx:
                raise ParseError(f"derived field {desc!r}: {name!r} is not a field")
            factors.append((idx[name], get_power(part)))
        return DerivedDesc(desc, 0, [(1.0, factors)])

    def _valid_name(self, name: str) -> bool:
        stripped = name.replace(" ", "")
        if stripped in ("", "LAP"):
            return True
        if len(stripped) < 3:
            raise ParseError(f"slice bounds out of range [:3] with length {len(stripped)}")
        if stripped[:3] == "LAP":
            stripped = stripped[3:]
        return self.is_brick_name(stripped) or self.is_field_name(stripped)

    def concrete_term(self, td: SubStringDelimiter) -> TermDesc:
        """ConcreteTerm (rhsBuilder.go:125-190) as data."""
        term = td.SubString
        sign = -1.0 if td.PreceedingDelimiter == "-" else 1.0
        brick_names: List[str] = []
        powers: List[float] = []
        for part in go_find_all(r"[^\*]*", term):
            name = go_find_string(r"^[^\^]*", part)
            if not self.is_field_name(name) and self.is_brick_name(name):
                brick_names.append(name)
                powers.append(get_power(part))
            elif not self._valid_name(name):
                raise ParseError(f"rhsBuilder: Name {name} is not defined!")
        fname = get_field_name(sort_factors(term), self.all_field_names())
        coef = complex(sign, 0.0)
        for bn, p in zip(brick_names, powers):
            coef = go_cmul(coef, go_cpow(self.scalars[bn], p))
        desc = TermDesc(KIND_CONCRETE, coef=coef, sign=sign, factors=list(zip(brick_names, powers)))
        if fname != "":
            desc.brick = self.brick_index(fname)
        if "LAP" in term:
            res = go_find_string(r"LAP*[^a-zA-Z]*", term)
            desc.has_lap = True
            desc.lap = int(get_power(res))
        return desc

    def _user_term(self, ut: UserTerm, prefixes: List[str]) -> TermDesc:
        """constructFunc (rhsBuilder.go:199-242): '-' flips the sign, LAP^n multiplies."""
        sign, lap = 1.0, 0
        for p in prefixes:
            if p == "-":
                sign = -sign
            elif p == "LAP^4":
                lap += 4
            elif p == "LAP^2":
                lap += 2
            elif p == "LAP":
                lap += 1
        if ut.kind == KIND_VOLCONS and not self.is_field_name(ut.params["indicator"]):
            raise ParseError("VolumeConservingLP: Indicator is not a derived field")
        table = self.user_tables[ut.kind]
        table.append(ut.params)
        return TermDesc(ut.kind, coef=complex(sign, 0.0), lap=lap, user=len(table) - 1)

    def build(self, eq: str) -> Tuple[List[TermDesc], List[TermDesc]]:
        """Build (rhsBuilder.go:26-54)."""
        sides = eq.split("=")
        if len(sides) != 2:
            raise ParseError("build: equality sign can only occur once")
        fld = field_name_from_leibniz(sides[0])
        terms: List[TermDesc] = []
        denum: List[TermDesc] = []
        for t in split_on_many(sides[1], ["+", "-"]):
            name = remove_known_prefixes(t.SubString)
            prefixes = get_known_prefixes(t.SubString) + [t.PreceedingDelimiter]
            if name in self.implicit_terms:
                denum.append(self._user_term(self.implicit_terms[name], prefixes))
            elif name in self.explicit_terms:
                terms.append(self._user_term(self.explicit_terms[name], prefixes))
            elif is_bilinear(t.SubString, fld, self.all_field_names()):
                stripped = SubStringDelimiter(t.SubString.replace(fld, ""), t.PreceedingDelimiter)
                denum.append(self.concrete_term(stripped))
            else:
                terms.append(self.concrete_term(t))
        return terms, denum

    def init(self) -> None:
        """Model.Init (model.go:244-260)."""
        self.rhs = []
        self.user_tables = {k: [] for k in _USER_KINDS}
        for eq in self.equations:
            self.rhs.append(self.build(eq))

    def summary(self) -> List[Tuple[int, int]]:
        """(len(Terms), len(Denum)) per equation, as Model.Summarize prints."""
        return [(len(t), len(d)) for t, d in self.rhs]
