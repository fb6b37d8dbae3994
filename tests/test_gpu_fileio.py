"""Device snapshots in the reference's file formats against the oracle's encodings.

Float64IO files are compared byte for byte with the big-endian encoding of the field the
solver returns; Uint8IO files with RealPartAsUint8 of it (pf/util.go:108-117).
"""
import os

import numpy as np
import pytest

import gopf_amd as G
from gopf_amd import workloads as W
from oracle import fileio as O

pytestmark = pytest.mark.gpu


def _ch_solver(dims, precision, fused=True, seed=0):
    n = int(np.prod(dims))
    c0 = W.ch_initial(seed, n)
    m = G.Model()
    m.AddField("conc", c0)
    for k, v in W.ch_scalars(1.0, 1.0).items():
        m.AddScalar(k, v)
    m.AddEquation(W.CH_EQUATION)
    m.Init()
    return G.Solver(m, dims, 0.1, precision=precision, fused=fused)


def _two_field_solver(N=16):
    rng = np.random.default_rng(2)
    m = G.Model()
    m.AddField("conc", 0.5 + 0.1 * rng.random(N * N))
    m.AddField("eta", 0.1 * rng.random(N * N))
    m.AddEquation("dconc/dt = LAP conc - conc*eta")
    m.AddEquation("deta/dt = -eta + LAP eta")
    m.Init()
    return G.Solver(m, [N, N], 0.05)


def _read(path):
    with open(path, "rb") as f:
        return f.read()


@pytest.mark.parametrize("case", ["fused_fp64", "fused_fp32", "general_fp64", "fused_3d_fp32"])
def test_float64io_bytes(tmp_path, case):
    if case == "fused_fp64":
        s = _ch_solver([64, 64], "fp64")
    elif case == "fused_fp32":
        s = _ch_solver([64, 64], "fp32")
    elif case == "fused_3d_fp32":
        s = _ch_solver([32, 32, 32], "fp32")
    else:
        s = _ch_solver([64, 64], "fp64", fused=False)
    s.Propagate(5)
    prefix = str(tmp_path / "snap")
    G.Float64IO(prefix).SaveFields(s, 3)
    s.FlushIO()
    assert _read(prefix + "_conc_3.bin") == O.float64_bytes(s.Field("conc").real)


def test_float64io_two_fields_and_xdmf(tmp_path):
    s = _two_field_solver()
    prefix = str(tmp_path / "run")
    w = G.Float64IO(prefix)
    seen = []
    s.Solve(3, 4, callbacks=[w.SaveFields, lambda sol, ep: seen.append({n: sol.Field(n).real for n in ("conc", "eta")})])
    for ep in range(3):
        for name in ("conc", "eta"):
            assert _read(f"{prefix}_{name}_{ep}.bin") == O.float64_bytes(seen[ep][name])
    G.WriteXDMF(str(tmp_path / "run.xdmf"), ["conc", "eta"], prefix, 3, [16, 16])
    text = open(tmp_path / "run.xdmf").read()
    for ep in range(3):
        assert f"{prefix}_conc_{ep}.bin" in text and os.path.exists(f"{prefix}_eta_{ep}.bin")


def test_snapshots_overlap_stepping(tmp_path):
    """Many snapshots in flight while the solver steps: every file holds its epoch's field."""
    s = _ch_solver([128, 128], "fp32")
    prefix = str(tmp_path / "ov")
    w = G.Float64IO(prefix)
    fields = []
    for ep in range(6):
        s.Propagate(20)
        w.SaveFields(s, ep)
        fields.append(s.Field("conc").real.copy())
    s.FlushIO()
    for ep in range(6):
        assert np.array_equal(G.LoadFloat64(f"{prefix}_conc_{ep}.bin"), fields[ep])


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_uint8io_bytes(tmp_path, precision):
    s = _ch_solver([64, 64], precision)
    s.Propagate(10)
    prefix = str(tmp_path / "u8")
    G.Uint8IO(prefix).SaveFields(s, 0)
    s.FlushIO()
    got = np.frombuffer(_read(prefix + "_conc_0.bin"), dtype=np.uint8)
    assert np.array_equal(got, O.real_part_as_uint8(s.Field("conc").real))


def test_csvio_device(tmp_path):
    s = _two_field_solver(8)
    s.Propagate(2)
    prefix = str(tmp_path / "csv")
    G.CsvIO(prefix, [8, 8]).SaveFields(s, 4)
    loaded = G.LoadCSV(prefix + "_4.csv")
    assert [n for n, _ in loaded] == ["conc", "eta"]
    for name, vals in loaded:
        assert np.allclose(vals.real, s.Field(name).real, atol=5e-7)


def test_fielddb_with_device_solver(tmp_path):
    s = _two_field_solver(8)
    db = G.FieldDB(str(tmp_path / "fdb.db"), [8, 8])
    s.Solve(2, 3, callbacks=[db.SaveFields])
    last = dict(db.LoadLast(db.simID))
    assert np.array_equal(last["conc"].real, s.Field("conc").real)
    assert np.array_equal(last["eta"].real, s.Field("eta").real)


def test_bad_directory_reports_at_flush(tmp_path):
    s = _ch_solver([16, 16], "fp64")
    G.Float64IO(str(tmp_path / "missing" / "x")).SaveFields(s, 0)
    with pytest.raises(G.GpfError):
        s.FlushIO()
    s.FlushIO()   # the error is reported once
