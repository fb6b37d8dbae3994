"""The product's C++ parser (through the C ABI) against the oracle's restatement.

Both restate pf/rhsBuilder.go:26-242 independently; they must agree term by
term (kind, coefficient, brick, LAP power) and on the derived fields, and both
must reproduce the reference's own parser tests.
"""
import numpy as np
import pytest

import gopf_amd as G
from oracle import parser as P

CORPUS = [
    # (fields, scalars, equations)
    (["conc"], {"m2": 2, "m6": 6, "m4": 4, "mk": 1},
     ["dconc/dt = m2*LAP conc - m6*LAP conc^2 + m4*LAP conc^3 - mk*LAP^2 conc"]),
    (["conc"], {"gamma": 2.0, "m1": -1.0},
     ["dconc/dt = LAP conc^3 + m1*LAP conc + m1*gamma*LAP^2 conc"]),          # examples/cahnHilliard
    (["concA", "concB", "concC"], {"kf": 2.0, "kr": 0.2},
     ["dconcA/dt = LAP concA - kf*concA^2*concB^3 + kr*concC",
      "dconcB/dt = LAP concB - kf*concA^2*concB^3 + kr*concC",
      "dconcC/dt = LAP concC - kr*concC + kf*concA^2*concB^3"]),
    (["cluster", "solute"], {}, ["dcluster/dt = LAP cluster", "dsolute/dt = solute*cluster"]),
    (["field"], {"rate": -1.0}, ["dfield/dt = rate*field"]),
    (["field"], {"rate": -1.0}, ["dfield/dt = field + rate*field^2"]),
    (["conc", "conc2"], {}, ["dconc/dt = conc2", "dconc2/dt = -conc2"]),
    (["field1", "field2"], {"ZERO": 0.0}, ["dfield1/dt = ZERO*field1", "dfield2/dt = ZERO*field1"]),
    (["current", "voltage", "magnetic"], {"resistance": 2.0},
     ["dcurrent/dt = voltage^2", "dvoltage/dt = resistance*current^2", "dmagnetic/dt = current*magnetic^3"]),
    (["conc"], {"a": 0.5, "b": -3.0}, ["dconc/dt = -a*LAP^2 conc + b*a^2*LAP^4 conc^2 - LAP conc"]),
    (["conc"], {"neg": -2.0}, ["dconc/dt = neg^3*LAP conc^2 - conc*conc"]),
    (["height"], {"eps": 0.3}, ["dheight/dt = eps*LAP height - height^3 + height^2.5"]),
]


def build_pair(fields, scalars, eqs):
    pm = G.Model()
    om = P.Model()
    for f in fields:
        pm.AddField(f, n=8)
        om.add_field(f, 8)
    for s, v in scalars.items():
        pm.AddScalar(s, v)
        om.add_scalar(s, v)
    for e in eqs:
        pm.AddEquation(e)
        om.add_equation(e)
    pm.Init()
    om.init()
    return pm, om


@pytest.mark.parametrize("case", range(len(CORPUS)))
def test_parsers_agree(case):
    pm, om = build_pair(*CORPUS[case])
    assert pm.AllFieldNames() == om.all_field_names()
    for eq in range(len(om.rhs)):
        pt, pd = pm.RHS(eq)
        ot, od = om.rhs[eq]
        for plist, olist in ((pt, ot), (pd, od)):
            assert len(plist) == len(olist)
            for a, b in zip(plist, olist):
                assert a.kind == b.kind and a.brick == b.brick and bool(a.has_lap) == b.has_lap
                assert a.lap == b.lap
                assert complex(a.coef_re, a.coef_im) == b.coef  # same formula, same rounding


def test_reference_reaction_diffusion_counts():
    """TestReactionDiffusion (model_test.go:45-108) through the product."""
    pm, _ = build_pair(*CORPUS[2])
    assert sorted(pm.AllFieldNames()) == sorted(["concA", "concB", "concC", "concA^2*concB^3"])
    counts = [tuple(len(x) for x in pm.RHS(i)) for i in range(3)]
    assert counts == [(2, 1), (2, 1), (1, 2)]


def test_reference_term_diffusion():
    """TestTermDiffusion (model_test.go:17-43)."""
    pm, _ = build_pair(["conc"], {}, ["dconc/dt = LAP conc"])
    t, d = pm.RHS(0)
    assert len(t) == 0 and len(d) == 1 and d[0].lap == 1 and d[0].brick == -1


@pytest.mark.parametrize("eq,ok", [("dconc/dt = LAP conc", True), ("dconc/dt = m1*conc", False),
                                   ("dconc/dt = LAP otherField", False), ("dc/dq = LAP conc", False),
                                   ("dconc/dt = a = b", False)])
def test_panics_become_errors(eq, ok):
    """TestPanicOnUnknownName / TestNameFromLeibniz: the reference panics, the ABI returns an error."""
    pm = G.Model()
    pm.AddField("conc", n=4)
    if ok:
        pm.AddEquation(eq)
        pm.Init()
    else:
        with pytest.raises(G.GpfError):
            pm.AddEquation(eq)
            pm.Init()


def test_reserved_prefix_in_name_rejected():
    """panicOnPrefixInName (model.go:480-485)."""
    pm = G.Model()
    with pytest.raises(G.GpfError):
        pm.AddField("LAPfield", n=4)
    with pytest.raises(G.GpfError):
        pm.AddScalar("-x", 1.0)


def test_user_term_prefixes():
    """TestNegativeSignBeforeUserDefined / constructFunc prefixes through the product."""
    pm = G.Model()
    pm.AddField("conc", n=4)
    pm.RegisterExplicitTerm("TERM", G.SquaredGradient("conc"))
    pm.AddEquation("dconc/dt = -TERM + LAP^2 TERM - -LAP TERM")
    pm.Init()
    t, d = pm.RHS(0)
    assert [(x.kind, x.coef_re, x.lap) for x in t] == [(1, -1.0, 0), (1, 1.0, 2), (1, -1.0, 1)]
    om = P.Model()
    om.add_field("conc", 4)
    om.register_explicit_term("TERM", P.KIND_SQGRAD, field="conc")
    om.add_equation("dconc/dt = -TERM + LAP^2 TERM - -LAP TERM")
    om.init()
    assert [(x.kind, x.coef.real, x.lap) for x in om.rhs[0][0]] == [(x.kind, x.coef_re, x.lap) for x in t]
