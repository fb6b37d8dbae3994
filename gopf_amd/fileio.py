"""Field output in gopf's file formats (pf/fileIO.go, pf/xdmf.go, pf/database.go).

The binary writers are solver callbacks like the reference's (``SaveFields(solver, epoch)``).
``Float64IO`` and ``Uint8IO`` snapshot the resident fields on the device: the file bytes
are packed on the solver's stream and written by a host thread while the solver keeps
stepping; ``Solver.FlushIO()`` (also run at the end of ``Solver.Solve`` and when the
solver is destroyed) waits for the files.  ``FieldDB`` keeps the reference's SQLite schema
and batches each field's inserts into one statement.
"""
from __future__ import annotations

import csv
import ctypes as C
import random
import sqlite3
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import GPF_IO_FLOAT64, GPF_IO_UINT8, _check, _dptr, _ints, lib


def _names(names: Sequence[str]):
    arr = (C.c_char_p * max(len(names), 1))(*[str(n).encode() for n in names])
    return arr


class Float64IO:
    """pf.Float64IO (fileIO.go:45-62): <Prefix>_<field>_<epoch>.bin, big-endian float64."""

    def __init__(self, prefix: str):
        self.Prefix = prefix

    def SaveFields(self, solver, epoch: int) -> None:
        _check(lib().gpf_solver_save_fields(solver._h, self.Prefix.encode(), int(epoch), GPF_IO_FLOAT64))


class Uint8IO:
    """pf.Uint8IO (fileIO.go:15-42): the real part quantised to uint8 between its min and max."""

    def __init__(self, prefix: str):
        self.Prefix = prefix

    def SaveFields(self, solver, epoch: int) -> None:
        _check(lib().gpf_solver_save_fields(solver._h, self.Prefix.encode(), int(epoch), GPF_IO_UINT8))


def NewFloat64IO(prefix: str) -> Float64IO:
    return Float64IO(prefix)


def NewUint8IO(prefix: str) -> Uint8IO:
    return Uint8IO(prefix)


class CsvIO:
    """pf.CsvIO (fileIO.go:97-115): <Prefix>_<epoch>.csv with X, Y, Z and one column per field."""

    def __init__(self, prefix: str, domain_size: Sequence[int]):
        self.Prefix = prefix
        self.DomainSize = list(domain_size)

    def SaveFields(self, solver, epoch: int) -> None:
        _check(lib().gpf_solver_save_csv(solver._h, self.Prefix.encode(), int(epoch)))


def SaveFloat64(fname: str, data) -> None:
    """SaveFloat64 (fileIO.go:87-95)."""
    a = np.ascontiguousarray(np.asarray(data, dtype=np.float64).reshape(-1))
    _check(lib().gpf_save_float64(fname.encode(), _dptr(a), a.size))


def LoadFloat64(fname: str) -> np.ndarray:
    """LoadFloat64 (fileIO.go:64-84)."""
    n = C.c_long()
    _check(lib().gpf_load_float64(fname.encode(), None, 0, C.byref(n)))
    out = np.zeros(n.value)
    _check(lib().gpf_load_float64(fname.encode(), _dptr(out), n.value, C.byref(n)))
    return out


def SaveCsv(fname: str, data: Sequence[Tuple[str, Sequence[float]]], domain_size: Sequence[int]) -> None:
    """SaveCsv (fileIO.go:147-200); data holds (name, values) pairs (CsvData)."""
    names = [d[0] for d in data]
    vals = np.ascontiguousarray(np.array([np.asarray(d[1]).real for d in data], dtype=np.float64).reshape(-1))
    n = len(data[0][1]) if data else 0
    dims = list(domain_size) if domain_size is not None else []
    _check(lib().gpf_save_csv(fname.encode(), _names(names), len(names), _dptr(vals) if vals.size else None, n,
                              _ints(dims) if dims else None, len(dims)))


def LoadCSV(fname: str) -> List[Tuple[str, np.ndarray]]:
    """LoadCSV (fileIO.go:117-145): the field columns after X, Y, Z as complex arrays."""
    with open(fname, newline="") as f:
        rows = list(csv.reader(f))
    header = rows[0]
    out = []
    for j, name in enumerate(header[3:]):
        vals = np.array([float(r[j + 3]) for r in rows[1:]], dtype=np.float64)
        out.append((name, vals.astype(np.complex128)))
    return out


def CreateXDMF(field_names: Sequence[str], prefix: str, num: int, domain_size: Sequence[int],
               indent: Optional[str] = "    ") -> str:
    """CreateXDMF (xdmf.go:83-160) encoded by xml.Encoder; indent None or "" is a bare encoder."""
    dims = list(domain_size)
    ln = C.c_long()
    ind = (indent or "").encode()
    _check(lib().gpf_create_xdmf(_names(field_names), len(field_names), prefix.encode(), int(num), _ints(dims),
                                 len(dims), ind, None, 0, C.byref(ln)))
    buf = C.create_string_buffer(ln.value + 1)
    _check(lib().gpf_create_xdmf(_names(field_names), len(field_names), prefix.encode(), int(num), _ints(dims),
                                 len(dims), ind, buf, ln.value + 1, C.byref(ln)))
    return buf.raw[:ln.value].decode()


def WriteXDMF(xdmf_file: str, fields: Sequence[str], prefix: str, num: int, domain_size: Sequence[int]) -> None:
    """WriteXDMF (xdmf.go:163-173)."""
    dims = list(domain_size)
    _check(lib().gpf_write_xdmf(xdmf_file.encode(), _names(fields), len(fields), prefix.encode(), int(num),
                                _ints(dims), len(dims)))


def _pos(domain_size: Sequence[int], n: int) -> np.ndarray:
    """pfutil.Pos (indexPositionConversion.go:24-44) for every node: rows of (X, Y, Z)."""
    i = np.arange(n, dtype=np.int64)
    if len(domain_size) == 2:
        return np.stack([i // domain_size[1], i % domain_size[1], np.zeros_like(i)], axis=1)
    if len(domain_size) == 3:
        d0, d1 = domain_size[0], domain_size[1]
        return np.stack([(i // d1) % d0, i % d1, i // (d0 * d1)], axis=1)
    raise ValueError("util: Domain size has to be either 2 or 3")


class FieldDB:
    """pf.FieldDB (database.go:14-80): fields, positions, attributes, comments and time series
    of a simulation in SQLite.  The schema and the stored rows are the reference's; each
    table's inserts go in one executemany inside one transaction (the reference prepares one
    statement per row, database.go:166-226)."""

    def __init__(self, db, domain_size: Optional[Sequence[int]] = None):
        self.DB = sqlite3.connect(db) if isinstance(db, str) else db
        self.DomainSize = list(domain_size) if domain_size is not None else None
        self.simID = 0
        self.initialized = False

    def initialize(self) -> None:
        """initialize (database.go:84-156)."""
        c = self.DB
        c.execute("CREATE TABLE IF NOT EXISTS simIDs (simID INTEGER UNIQUE, creationTime TEXT)")
        c.execute("CREATE TABLE IF NOT EXISTS positions (id INTEGER PRIMARY KEY, X INTEGER, Y INTEGER, Z INTEGER)")
        c.execute("CREATE TABLE IF NOT EXISTS fields (id INTEGER PRIMARY KEY, name TEXT, "
                  "value REAL, positionId INTEGER, timestep INTEGER, simID INTEGER, "
                  "FOREIGN KEY(positionId) REFERENCES positions(id), "
                  "FOREIGN KEY(simID) REFERENCES simIDs(simID))")
        c.execute("CREATE TABLE IF NOT EXISTS simAttributes (key TEXT, value REAL, simID INTEGER, "
                  "FOREIGN KEY(simID) REFERENCES fields(simID))")
        c.execute("CREATE TABLE IF NOT EXISTS comments (simID INTEGER, value TEXT, "
                  "FOREIGN KEY(simID) REFERENCES simIds(simID))")
        c.execute("CREATE TABLE IF NOT EXISTS simTextAttributes (key TEXT, value TEXT, "
                  "simID INTEGER, FOREIGN KEY(simID) REFERENCES simIds(simID))")
        c.execute("CREATE TABLE IF NOT EXISTS timeseries (key TEXT, value REAL, "
                  "timestep INTEGER, simID INTEGER, FOREIGN KEY(simID) REFERENCES simIds(simID))")
        if self.simID == 0:
            # rand.Int31 of a time-seeded source: a non-negative 31-bit id
            while self.simID == 0:
                self.simID = random.SystemRandom().randrange(0, 2 ** 31)
            c.execute("INSERT INTO simIDs (simID, creationTime) VALUES (?, datetime('now', 'localtime'))",
                      (self.simID,))
        c.commit()
        self.initialized = True

    def positionTableIsPopulated(self) -> bool:
        """positionTableIsPopulated (database.go:158-170)."""
        (n,) = self.DB.execute("SELECT COUNT(*) FROM positions").fetchone()
        return n > 1

    def populatePositionsTable(self) -> None:
        """populatePositionsTable (database.go:172-197)."""
        n = int(np.prod(self.DomainSize))
        pos = _pos(self.DomainSize, n)
        with self.DB:
            self.DB.executemany("INSERT INTO positions (X, Y, Z) VALUES (?, ?, ?)", map(tuple, pos.tolist()))

    def domainSizeOk(self) -> bool:
        """domainSizeOk (database.go:344-371)."""
        if self.DomainSize is None:
            return True
        for i, col in enumerate(["X", "Y", "Z"][:len(self.DomainSize)]):
            (mx,) = self.DB.execute(f"SELECT MAX({col}) FROM positions").fetchone()
            if (mx if mx is not None else 0) != self.DomainSize[i] - 1:
                return False
        return True

    def insertRealPart(self, name: str, timestep: int, values) -> None:
        """insertRealPart (database.go:200-226)."""
        if not self.domainSizeOk():
            raise RuntimeError("FieldDB: Domain size does not match the positions table in the database")
        v = np.asarray(values)
        if self.DomainSize is not None and v.size != int(np.prod(self.DomainSize)):
            raise RuntimeError("FieldDB: The passed array does not match the specified domain size")
        re = np.real(v).astype(np.float64).tolist()
        sid = self.simID
        with self.DB:
            self.DB.executemany("INSERT INTO fields (name, value, positionId, timestep, simID) VALUES (?, ?, ?, ?, ?)",
                                ((name, re[i], i, int(timestep), sid) for i in range(len(re))))

    def SaveFields(self, solver, epoch: int) -> None:
        """SaveFields (database.go:230-244): a solver callback."""
        if not self.initialized:
            self.initialize()
        if not self.positionTableIsPopulated():
            self.populatePositionsTable()
        for name in solver.model.FieldNames():
            self.insertRealPart(name, epoch, solver.Field(name))

    def Comment(self, comment: str) -> None:
        """Comment (database.go:247-259)."""
        comment = comment.replace("\n", " ")
        if not self.initialized:
            self.initialize()
        with self.DB:
            self.DB.execute("INSERT INTO comments (simID, value) VALUES (?, ?)", (self.simID, comment))

    def SetAttr(self, attr: Dict[str, float]) -> None:
        """SetAttr (database.go:262-287)."""
        if not self.initialized:
            self.initialize()
        with self.DB:
            self.DB.executemany("INSERT INTO simAttributes (key, value, simID) VALUES (?, ?, ?)",
                                [(k, float(v), self.simID) for k, v in attr.items()])

    def SetTextAttr(self, attr: Dict[str, str]) -> None:
        """SetTextAttr (database.go:290-315)."""
        if not self.initialized:
            self.initialize()
        with self.DB:
            self.DB.executemany("INSERT INTO simTextAttributes (key, value, simID) VALUES (?, ?, ?)",
                                [(k, str(v), self.simID) for k, v in attr.items()])

    def TimeSeries(self, data: Dict[str, float], timestep: int) -> None:
        """TimeSeries (database.go:318-341)."""
        if not self.initialized:
            self.initialize()
        with self.DB:
            self.DB.executemany("INSERT INTO timeseries (key, value, timestep, simID) VALUES (?, ?, ?, ?)",
                                [(k, float(v), int(timestep), self.simID) for k, v in data.items()])

    def Load(self, simID: int, timestep: int) -> List[Tuple[str, np.ndarray]]:
        """Load (database.go:376-421): the fields of simID at timestep, ordered by name."""
        names = [r[0] for r in self.DB.execute(
            "SELECT DISTINCT name FROM fields WHERE simID=? AND timestep=? ORDER BY name", (simID, timestep))]
        (n,) = self.DB.execute("SELECT COUNT(*) FROM positions").fetchone()
        fields = {k: np.zeros(n, dtype=np.complex128) for k in names}
        for name, value, pid in self.DB.execute(
                "SELECT name, value, positionId FROM fields WHERE simID=? AND timestep=?", (simID, timestep)):
            fields[name][pid] = complex(value, 0.0)
        return [(k, fields[k]) for k in names]

    def LoadLast(self, simID: int) -> List[Tuple[str, np.ndarray]]:
        """LoadLast (database.go:425-436)."""
        (ts,) = self.DB.execute("SELECT MAX(timestep) FROM fields WHERE simID=?", (simID,)).fetchone()
        return self.Load(simID, ts if ts is not None else 0)


class BatchFieldDB:
    """Every sim of a ``Batch`` under its own FieldDB ``simID`` (database.go:67-80).

    The reference runs a sweep as one ``pf.Solver`` per sim, each saving through its own
    ``FieldDB`` and so under its own random simID.  A database keeps a single positions table
    for one domain size (database.go:159-197, 344-371), so the sims of one grid shape share a
    database file here, each with its own simID; ``path`` maps a shape (ny, nx) to that
    class's file (or is a format string with ``{ny}`` and ``{nx}``).  ``SaveFields(epoch)``
    stores every sim's field as ``FieldDB.SaveFields`` stores a solver's, in one transaction
    per sim; ``Attrs(sim, attrs)`` records a sim's parameters as ``SetAttr`` does."""

    def __init__(self, batch, path, field_name: str = "conc"):
        self.batch = batch
        self.field_name = field_name
        self.conns: Dict[Tuple[int, ...], sqlite3.Connection] = {}
        self.dbs: List[FieldDB] = []
        for s in range(len(batch)):
            shape = tuple(int(v) for v in batch.shapes[s])
            if shape not in self.conns:
                p = path(shape) if callable(path) else str(path).format(ny=shape[0], nx=shape[1])
                self.conns[shape] = sqlite3.connect(p)
            self.dbs.append(FieldDB(self.conns[shape], list(shape)))
        used = set()
        for db in self.dbs:
            db.initialize()
            # simIDs are random 31-bit integers; no two sims of one sweep share one
            while db.simID in used:
                with db.DB:
                    db.DB.execute("DELETE FROM simIDs WHERE simID=?", (db.simID,))
                db.simID = 0
                db.initialize()
            used.add(db.simID)

    def simID(self, sim: int) -> int:
        return self.dbs[sim].simID

    def Attrs(self, sim: int, attrs: Dict[str, float]) -> None:
        self.dbs[sim].SetAttr(attrs)

    def SaveFields(self, epoch: int) -> None:
        for s, db in enumerate(self.dbs):
            if not db.positionTableIsPopulated():
                db.populatePositionsTable()
            db.insertRealPart(self.field_name, epoch, self.batch.Field(s))

    def close(self) -> None:
        for c in self.conns.values():
            c.close()
