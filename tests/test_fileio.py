"""Field file formats on the host side (no GPU): the oracle pinned by the reference's
golden strings, and the library's writers checked against the oracle.

Golden strings: TestTopology, TestDataItem, TestGeometry (pf/xdmf_test.go:9-49) and the
MarshalIndent example of Go's encoding/xml documentation.  Known answers: TestSaveLoad,
TestSaveLoadCsv (pf/fileIO_test.go) and the FieldDB tests (pf/database_test.go), restated.
"""
import os
import sqlite3
import xml.etree.ElementTree as ET

import numpy as np
import pytest

import gopf_amd as G
from oracle import fileio as O


# ---------------------------------------------------------------------------
# the oracle's XML marshaller against golden strings
# ---------------------------------------------------------------------------

def test_oracle_topology_golden():
    el = O.topology(name="topo", type_="3DCoRectMesh", dimensions="100 500 500")
    assert O.marshal(el) == '<Topology name="topo" TopologyType="3DCoRectMesh" Dimensions="100 500 500"></Topology>'


def test_oracle_data_item_golden():
    el = O.data_item(dim=2, value="0.0 0.0", format_="XML")
    assert O.marshal(el) == '<DataItem Dimension="2" Format="XML">0.0 0.0</DataItem>'


def test_oracle_geometry_golden():
    el = O.geometry(name="geo", items=[O.data_item(), O.data_item()])
    assert O.marshal(el) == '<Geometry name="geo"><DataItem></DataItem><DataItem></DataItem></Geometry>'


def test_oracle_indent_rule_golden():
    """encoding/xml's ExampleMarshalIndent output (prefix dropped): text-only elements stay on
    one line, a parent's end tag gets its own line."""
    person = ("person", [("id", "13")], [
        ("name", [], [("first", [], [], "John"), ("last", [], [], "Doe")], ""),
        ("age", [], [], "42"), ("Married", [], [], "false"), ("City", [], [], "Hanga Roa"),
        ("State", [], [], "Easter Island")], "")
    expect = "\n".join([
        '<person id="13">',
        "    <name>",
        "        <first>John</first>",
        "        <last>Doe</last>",
        "    </name>",
        "    <age>42</age>",
        "    <Married>false</Married>",
        "    <City>Hanga Roa</City>",
        "    <State>Easter Island</State>",
        "</person>"])
    assert O.marshal(person, indent="    ") == expect


def test_oracle_planted_mistake_fails_a_golden():
    """Dropping omitempty (writing empty attributes) is caught by TestGeometry's golden."""
    el = O.geometry(name="geo", items=[O.data_item(), O.data_item()])
    wrong = O.marshal(el, keep=("TimeType", "Type"))
    assert wrong != '<Geometry name="geo"><DataItem></DataItem><DataItem></DataItem></Geometry>'


# ---------------------------------------------------------------------------
# the library against the oracle
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("dims", [[128, 128], [16, 16, 16]])
@pytest.mark.parametrize("indent", ["", "    "])
def test_create_xdmf_matches_oracle(dims, indent):
    names = ["conc", "eta"]
    got = G.CreateXDMF(names, "myprefix", 2, dims, indent=indent)
    assert got == O.marshal(O.create_xdmf(names, "myprefix", 2, dims), indent=indent)


def test_create_xdmf_escapes_and_many_steps():
    names = ["a&b", 'q"x']
    got = G.CreateXDMF(names, "dir/p<1>", 7, [8, 8])
    assert got == O.marshal(O.create_xdmf(names, "dir/p<1>", 7, [8, 8]), indent="    ")


def test_write_xdmf_parses(tmp_path):
    path = str(tmp_path / "out.xdmf")
    G.WriteXDMF(path, ["conc"], "run", 3, [32, 32, 32])
    root = ET.parse(path).getroot()
    assert root.tag == "Xdmf"
    series = root.find("Domain/Grid")
    assert series.get("CollectionType") == "Temporal"
    grids = series.findall("Grid")
    assert [g.get("Name") for g in grids] == ["T1", "T2", "T3"]
    items = [g.find("Attribute/DataItem") for g in grids]
    assert [i.text for i in items] == ["run_conc_0.bin", "run_conc_1.bin", "run_conc_2.bin"]
    assert all(i.get("Endian") == "Big" and i.get("Precision") == "8" and i.get("Dimensions") == "32 32 32"
               for i in items)
    assert series.find("Time/DataItem").text == "0.000000 0.333333 0.666667 3"


def test_xdmf_rejects_1d():
    with pytest.raises(G.GpfError):
        G.CreateXDMF(["c"], "p", 1, [8])


def test_save_load_float64(tmp_path):
    """TestSaveLoad (pf/fileIO_test.go:11-35), with the file bytes against the oracle."""
    data = np.arange(64, dtype=np.float64) * 1.25 - 7.0
    data[3] = -0.0
    data[5] = np.inf
    data[6] = np.nan
    path = str(tmp_path / "myfile_conc_0.bin")
    G.SaveFloat64(path, data)
    with open(path, "rb") as f:
        raw = f.read()
    assert raw == O.float64_bytes(data)
    back = G.LoadFloat64(path)
    assert back.shape == data.shape
    assert np.array_equal(back.view(np.uint64), data.view(np.uint64))


def test_save_load_csv(tmp_path):
    """TestSaveLoadCsv (pf/fileIO_test.go:37-69) and the first records' text."""
    N = 8
    i = np.arange(N * N, dtype=np.float64)
    f1, f2 = i - i * i, i
    path = str(tmp_path / "my_csv_file_0.csv")
    G.SaveCsv(path, [("field1", f1), ("field2", f2)], [N, N])
    loaded = G.LoadCSV(path)
    assert [n for n, _ in loaded] == ["field1", "field2"]
    assert np.allclose(loaded[0][1].real, f1, atol=1e-8) and np.allclose(loaded[1][1].real, f2, atol=1e-8)
    with open(path) as f:
        lines = f.read().split("\n")
    assert lines[0] == "X,Y,Z,field1,field2"
    assert lines[1] == "0,0,0,0.000000,0.000000"
    assert lines[9] == "1,0,0,-56.000000,8.000000"
    assert lines[-1] == ""


def test_save_csv_3d_positions(tmp_path):
    dims = [2, 3, 4]
    n = 24
    path = str(tmp_path / "c.csv")
    G.SaveCsv(path, [("v", np.arange(n) + 0.5)], dims)
    rows = [r.split(",") for r in open(path).read().strip().split("\n")[1:]]
    for k, r in enumerate(rows):
        # pfutil.Pos for 3D (indexPositionConversion.go:24-29): col = k % D1, row = (k / D1) % D0, depth = k / (D0 D1)
        assert [int(r[0]), int(r[1]), int(r[2])] == [(k // 3) % 2, k % 3, k // 6]
        assert r[3] == "%f" % (k + 0.5)


def test_save_csv_checks_domain():
    with pytest.raises(G.GpfError):
        G.SaveCsv("/tmp/never.csv", [("v", np.zeros(10))], [3, 3])


def test_uint8_oracle_closed_forms():
    x = np.array([2.0, -1.0, 0.5, 3.0, np.nan])
    u = O.real_part_as_uint8(x[:4])
    assert list(u) == [191, 0, 95, 255]   # 255 (x - min) / (max - min), truncated
    assert list(O.real_part_as_uint8(np.full(5, 0.7))) == [0] * 5   # max = min + 1 when flat
    # NaN elements are skipped by the min / max scan
    assert list(O.real_part_as_uint8(x)[:4]) == [191, 0, 95, 255]


# ---------------------------------------------------------------------------
# FieldDB (pf/database_test.go)
# ---------------------------------------------------------------------------

class _FakeModel:
    def __init__(self, names):
        self._names = names

    def FieldNames(self):
        return list(self._names)


class _FakeSolver:
    """Stands for a device solver: Field(name) returns the stored array."""

    def __init__(self, fields):
        self.model = _FakeModel(list(fields))
        self.fields = fields

    def Field(self, name):
        return np.asarray(self.fields[name], dtype=np.complex128)


def test_fielddb_initialize(tmp_path):
    db = G.FieldDB(str(tmp_path / "testinit.db"))
    db.initialize()
    assert db.initialized and db.simID != 0
    names = [r[0] for r in db.DB.execute("SELECT name FROM sqlite_master WHERE type='table' ORDER BY name")]
    assert names == ["comments", "fields", "positions", "simAttributes", "simIDs", "simTextAttributes", "timeseries"]
    assert not db.positionTableIsPopulated()


def test_fielddb_positions(tmp_path):
    db = G.FieldDB(str(tmp_path / "p.db"), [3, 3])
    db.initialize()
    db.populatePositionsTable()
    rows = db.DB.execute("SELECT X, Y, Z FROM positions ORDER BY id").fetchall()
    assert rows == [(0, 0, 0), (0, 1, 0), (0, 2, 0), (1, 0, 0), (1, 1, 0), (1, 2, 0), (2, 0, 0), (2, 1, 0), (2, 2, 0)]


def test_fielddb_insert_and_save(tmp_path):
    db = G.FieldDB(str(tmp_path / "f.db"), [3, 3])
    db.initialize()
    db.populatePositionsTable()
    db.insertRealPart("myfield", 2, np.arange(9, dtype=complex))
    rows = db.DB.execute("SELECT name, timestep, value, simID FROM fields ORDER BY id").fetchall()
    assert rows == [("myfield", 2, float(k), db.simID) for k in range(9)]
    db2 = G.FieldDB(str(tmp_path / "s.db"), [3, 3])
    db2.SaveFields(_FakeSolver({"field1": np.zeros(9), "field2": np.ones(9)}), 1)
    assert db2.DB.execute("SELECT COUNT(*) FROM positions").fetchone()[0] == 9
    assert db2.DB.execute("SELECT COUNT(*) FROM fields").fetchone()[0] == 18
    with pytest.raises(RuntimeError):
        db2.insertRealPart("x", 0, np.zeros(8))


def test_fielddb_comment_attrs_timeseries(tmp_path):
    db = G.FieldDB(str(tmp_path / "a.db"))
    db.Comment("This is\na comment")
    assert db.DB.execute("SELECT simID, value FROM comments").fetchall() == [(db.simID, "This is a comment")]
    db.SetAttr({"temperature": 20.0, "concentration": 0.8})
    assert db.DB.execute("SELECT key, value FROM simAttributes ORDER BY key").fetchall() == [
        ("concentration", 0.8), ("temperature", 20.0)]
    db.SetTextAttr({"name": "simulation1", "node": "sophus"})
    assert db.DB.execute("SELECT key, value FROM simTextAttributes ORDER BY key").fetchall() == [
        ("name", "simulation1"), ("node", "sophus")]
    db.TimeSeries({"energy": -0.1, "concentration": 0.5}, 0)
    db.TimeSeries({"energy": -0.1, "concentration": 0.5}, 1)
    rows = db.DB.execute("SELECT key, value, timestep FROM timeseries ORDER BY timestep ASC, key DESC").fetchall()
    assert rows == [("energy", -0.1, 0), ("concentration", 0.5, 0), ("energy", -0.1, 1), ("concentration", 0.5, 1)]


def test_fielddb_load(tmp_path):
    """TestLoadFields (pf/database_test.go:364-420) with a stand-in solver."""
    db = G.FieldDB(str(tmp_path / "l.db"), [4, 4])
    for step in range(10):
        s = _FakeSolver({"temperature": np.full(16, 0.1 * (step + 1)), "conc": np.full(16, -0.1 * (step + 1))})
        db.SaveFields(s, step)
    for step in range(10):
        f = db.Load(db.simID, step)
        assert [n for n, _ in f] == ["conc", "temperature"]
        assert np.allclose(f[0][1], -0.1 * (step + 1)) and np.allclose(f[1][1], 0.1 * (step + 1))
    last = db.LoadLast(db.simID)
    assert np.allclose(last[1][1], 1.0)


class _FakeBatch:
    """A Batch's host interface (shapes, len, Field) over fixed arrays, so that the FieldDB
    mapping runs without a device."""

    def __init__(self, classes):
        self.shapes, self.data = [], []
        rng = np.random.default_rng(0)
        for shape, count in classes:
            for _ in range(count):
                self.shapes.append(tuple(shape))
                self.data.append(rng.random(int(np.prod(shape))).astype(np.complex128))

    def __len__(self):
        return len(self.shapes)

    def Field(self, s):
        return self.data[s]


def test_batch_fielddb_gives_every_sim_its_own_simid(tmp_path):
    """A sweep stored as the reference stores one solver per sim (database.go:67-80,
    230-244): every sim under its own simID, sims of one shape in one database (one positions
    table per domain size), each sim's field loadable by its simID (Load, database.go:376-421)."""
    b = _FakeBatch([((4, 6), 3), ((8, 4), 2)])
    bdb = G.BatchFieldDB(b, str(tmp_path / "sweep_{ny}x{nx}.db"))
    for s in range(len(b)):
        bdb.Attrs(s, {"M": 1.0 + s, "kappa": 0.5 * s})
    bdb.SaveFields(7)
    ids = [bdb.simID(s) for s in range(len(b))]
    assert len(set(ids)) == len(b) and all(0 < i < 2 ** 31 for i in ids)
    for s in range(len(b)):
        db = G.FieldDB(bdb.conns[b.shapes[s]], list(b.shapes[s]))
        (name, field), = db.Load(ids[s], 7)
        assert name == "conc" and np.array_equal(field.real, b.data[s].real)
        (m,) = db.DB.execute("SELECT value FROM simAttributes WHERE simID=? AND key='M'", (ids[s],)).fetchone()
        assert m == 1.0 + s
    # one positions table per shape, each holding its grid
    for shape, conn in bdb.conns.items():
        (n,) = conn.execute("SELECT COUNT(*) FROM positions").fetchone()
        assert n == int(np.prod(shape))
    assert len(bdb.conns) == 2
    bdb.close()
