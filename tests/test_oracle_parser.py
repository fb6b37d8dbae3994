"""The reference's parser tests, restated against the oracle parser.

Sources: pf/rhsBuilder_test.go, pf/model_test.go, pf/util_test.go.
"""
import math

import pytest

from oracle import parser as P


def test_name_from_leibniz():
    assert P.field_name_from_leibniz("dc/dt") == "c"
    assert P.field_name_from_leibniz("dkappa/dt") == "kappa"
    for bad in ["dc", "ac/dt", "dc/dq"]:
        with pytest.raises(P.ParseError):
            P.field_name_from_leibniz(bad)


@pytest.mark.parametrize("field,expr,expect,allf", [
    ("c", "2*c", True, ["c"]),
    ("conc", "conc^2", False, ["conc"]),
    ("c", "c*n*r", False, ["c", "n", "r"]),
    ("voltage", "voltage^1.62", False, ["voltage"]),
    ("voltage", "current*voltage^1.0", False, ["voltage", "current"]),
    ("current", "P*current^-2", False, ["current"]),
])
def test_is_bilinear(field, expr, expect, allf):
    assert P.is_bilinear(expr, field, allf) is expect


def _concrete_model():
    m = P.Model()
    for name in ("current", "voltage", "magnetic"):
        m.add_field(name, 1)
    m.add_scalar("resistance", 2.0)
    m.add_equation("dcurrent/dt = voltage^2")
    m.add_equation("dvoltage/dt = resistance*current^2")
    m.add_equation("dmagnetic/dt = current*magnetic^3")
    return m


def test_concrete_term_structure():
    """TestConcreteTerm (rhsBuilder_test.go:95-157): values 8, 16, -8, -16 at the single node."""
    m = _concrete_model()
    vals = {"current": 2.0, "voltage": -4.0, "magnetic": 1.5}
    names = m.all_field_names()
    for expr, sign, expect in [("resistance*current^2", "+", 8.0), ("voltage^2", "+", 16.0),
                               ("resistance*current^2", "-", -8.0), ("voltage^2", "-", -16.0)]:
        t = m.concrete_term(P.SubStringDelimiter(expr, sign))
        # evaluate the derived brick with cmplx.Pow semantics
        d = m.derived[t.brick - len(m.fields)]
        v = 1 + 0j
        for f, p in d.terms[0][1]:
            v = P.go_cmul(v, P.go_cpow(complex(vals[m.fields[f][0]]), p))
        got = P.go_cmul(t.coef, v)
        assert abs(got - expect) < 1e-10, (expr, sign, got)
        assert not t.has_lap
    assert sorted(names) == sorted(["current", "voltage", "magnetic", "voltage^2", "current^2",
                                    "current*magnetic^3"])


@pytest.mark.parametrize("expr,should_panic", [
    ("LAP conc", False), ("conc", False), ("m1*conc", True), ("m1*LAP conc", True),
    ("LAP otherField", True)])
def test_panic_on_unknown_name(expr, should_panic):
    m = P.Model()
    m.add_field("conc", 8)
    if should_panic:
        with pytest.raises(P.ParseError):
            m.concrete_term(P.SubStringDelimiter(expr.replace(" ", ""), ""))
    else:
        m.concrete_term(P.SubStringDelimiter(expr.replace(" ", ""), ""))


@pytest.mark.parametrize("s,expect", [
    ("mystring", "mystring"), ("*mystring", "mystring"), ("LAPmystring", "mystring"),
    ("LAP*mystring", "mystring"), ("*LAP*mystring", "mystring"), ("*LAPLAPmystring", "mystring"),
    ("*LAP^2mystring", "mystring")])
def test_remove_known_prefixes(s, expect):
    assert P.remove_known_prefixes(s) == expect


@pytest.mark.parametrize("s,expect", [
    ("mystring", []), ("*mystring", ["*"]), ("LAPmystring", ["LAP"]), ("LAP*mystring", ["LAP", "*"]),
    ("*LAP*mystring", ["*", "LAP", "*"]), ("*LAPLAPmystring", ["*", "LAP", "LAP"]),
    ("*LAP^2mystring", ["*", "LAP^2"])])
def test_get_known_prefixes(s, expect):
    assert P.get_known_prefixes(s) == expect


def test_term_diffusion():
    """TestTermDiffusion (model_test.go:17-43)."""
    m = P.Model()
    m.add_field("conc", 2)
    m.add_equation("dconc/dt = LAP conc")
    m.init()
    terms, denum = m.rhs[0]
    assert len(terms) == 0 and len(denum) == 1
    t = denum[0]
    assert t.brick == -1 and t.has_lap and t.lap == 1 and t.coef == 1


def test_reaction_diffusion():
    """TestReactionDiffusion (model_test.go:45-108)."""
    m = P.Model()
    for n in ("concA", "concB", "concC"):
        m.add_field(n, 2)
    m.add_scalar("kf", 2.0)
    m.add_scalar("kr", 0.2)
    m.add_equation("dconcA/dt = LAP concA - kf*concA^2*concB^3 + kr*concC")
    m.add_equation("dconcB/dt = LAP concB - kf*concA^2*concB^3 + kr*concC")
    m.add_equation("dconcC/dt = LAP concC - kr*concC + kf*concA^2*concB^3")
    m.init()
    assert sorted(m.all_field_names()) == sorted(["concA", "concB", "concC", "concA^2*concB^3"])
    assert m.summary() == [(2, 1), (2, 1), (1, 2)]


def test_non_alphabetic_order():
    """TestNonAlphabeticOrder (rhsBuilder_test.go:249-272): 'solute*cluster' -> derived 'cluster*solute'."""
    m = P.Model()
    m.add_field("cluster", 8)
    m.add_field("solute", 8)
    m.add_equation("dcluster/dt = LAP cluster")
    m.add_equation("dsolute/dt = solute*cluster")
    m.init()
    terms, denum = m.rhs[1]
    assert len(terms) == 1 and len(denum) == 0
    assert m.all_field_names()[terms[0].brick] == "cluster*solute"


def test_split_on_many():
    """TestSplitOnMany (util_test.go)."""
    got = P.split_on_many("a+b", ["+"])
    assert [s.SubString for s in got] == ["a", "b"]
    got = P.split_on_many("a+b+cd", ["+"])
    assert [s.SubString for s in got] == ["a", "b", "cd"]
    got = P.split_on_many("a+b-cd", ["+", "-"])
    assert [s.SubString for s in got] == ["a", "b", "cd"]
    got = P.split_on_many("cdb-the+two", ["+", "-", "7"])
    assert sorted(s.SubString for s in got) == ["cdb", "the", "two"]  # the reference test sorts
    assert [s.SubString for s in got] == ["two", "cdb", "the"]
    # breadth-first queue order and inherited delimiters (util.go:166-198)
    got = P.split_on_many("-a+b-c", ["+", "-"])
    assert [(s.SubString, s.PreceedingDelimiter) for s in got] == [("a", "-"), ("b", "+"), ("c", "-")]


def test_get_field_name():
    """TestGetFieldName (util_test.go)."""
    names = ["conc1", "conc2", "conc3", "conc1^2*conc2", "conc3^3", "conc2^4*conc1^2", "conc1^2"]
    assert P.get_field_name("conc1^2*conc2*otherstuff", names) == "conc1^2*conc2"
    assert P.get_field_name("*randomstuff*conc1^2*otherstuff", names) == "conc1^2"


def test_constructor_prefix_handling():
    """TestConstructorWithPrefixHandling (rhsBuilder_test.go:440-509) on a user term."""
    m = P.Model()
    m.add_field("conc", 4)
    m.register_explicit_term("TERM", P.KIND_SQGRAD, field="conc")
    cases = [(["-"], -1, 0), (["-", "-"], 1, 0), (["-", "-", "-"], -1, 0), (["LAP"], 1, 1),
             (["LAP", "-"], -1, 1), (["-", "LAP"], -1, 1), (["LAP^2"], 1, 2), (["LAP^4"], 1, 4),
             (["LAP^4", "-", "-"], 1, 4), (["-", "LAP^4", "-"], 1, 4)]
    for prefixes, sign, lap in cases:
        t = m._user_term(m.explicit_terms["TERM"], prefixes)
        assert t.coef.real == sign and t.lap == lap


def test_negative_sign_before_user_defined():
    """TestNegativeSignBeforeUserDefined (rhsBuilder_test.go:296-360), explicit and implicit."""
    m = P.Model()
    m.add_field("conc", 8)
    m.register_explicit_term("TERM", P.KIND_SQGRAD, field="conc")
    terms, denum = m.build("dconc/dt=-TERM")
    assert len(terms) == 1 and len(denum) == 0 and terms[0].coef == -1
    m = P.Model()
    m.add_field("conc", 8)
    m.register_implicit_term("TERM", P.KIND_SQGRAD, field="conc")
    terms, denum = m.build("dconc/dt=-TERM")
    assert len(terms) == 0 and len(denum) == 1 and denum[0].coef == -1


def test_cahn_hilliard_classification():
    """The CH string of examples/cahnHilliard/main.go:33 form: LAP terms implicit, powers explicit."""
    m = P.Model()
    m.add_field("conc", 16)
    for s, v in (("m2", 2.0), ("m6", 6.0), ("m4", 4.0), ("mk", 1.0)):
        m.add_scalar(s, v)
    m.add_equation("dconc/dt = m2*LAP conc - m6*LAP conc^2 + m4*LAP conc^3 - mk*LAP^2 conc")
    m.init()
    terms, denum = m.rhs[0]
    names = m.all_field_names()
    assert names == ["conc", "conc^2", "conc^3"]
    assert [(names[t.brick], t.coef.real, t.lap) for t in terms] == [("conc^2", -6.0, 1), ("conc^3", 4.0, 1)]
    assert [(t.brick, t.coef.real, t.lap) for t in denum] == [(-1, 2.0, 1), (-1, -1.0, 2)]


def test_scalar_negative_value_pow_rounding():
    """Go's cmplx.Pow(-1, 1) carries sin(pi) in the imaginary part (math/cmplx/pow.go)."""
    v = P.go_cpow(-1 + 0j, 1.0)
    assert v.real == -1.0 and abs(v.imag - math.sin(math.pi)) == 0.0
