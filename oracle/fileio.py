"""File formats of pf/fileIO.go and pf/xdmf.go (TEST INFRASTRUCTURE ONLY).

* ``marshal`` restates what Go's encoding/xml Encoder writes for a tree of structs
  (start tags with the non-empty omitempty attributes in field order, character data
  escaped with EscapeText, no self-closing tags, and the printer's writeIndent rule when an
  indent is set).  It is pinned by the reference's golden strings TestTopology,
  TestDataItem and TestGeometry (pf/xdmf_test.go:9-49).
* ``create_xdmf`` builds CreateXDMF's tree (pf/xdmf.go:83-160) from the struct
  definitions (xdmf.go:9-80) and marshals it.
* ``float64_bytes`` is SaveFloat64's encoding (fileIO.go:87-95): big-endian IEEE 754.
* ``real_part_as_uint8`` is RealPartAsUint8 (util.go:108-117) with MinReal / MaxReal
  (pfutil/sliceOperations.go:61-80) and Go's truncating float-to-integer conversion.
"""
from __future__ import annotations

from typing import List, Optional, Sequence, Tuple

import numpy as np

# an element: (name, [(attr, value)], children, chardata); empty attribute values are
# dropped (omitempty) except those listed in keep
Element = Tuple[str, List[Tuple[str, str]], list, str]


def _escape(s: str) -> str:
    out = []
    for ch in s:
        out.append({'"': "&#34;", "'": "&#39;", "&": "&amp;", "<": "&lt;", ">": "&gt;",
                    "\t": "&#x9;", "\n": "&#xA;", "\r": "&#xD;"}.get(ch, ch))
    return "".join(out)


def marshal(el: Element, indent: str = "", keep: Sequence[str] = ("TimeType",)) -> str:
    state = {"depth": 0, "indented_in": False, "put_newline": False}
    out: List[str] = []

    def write_indent(delta: int) -> None:
        if not indent:
            return
        if delta < 0:
            state["depth"] -= 1
            if state["indented_in"]:
                state["indented_in"] = False
                return
            state["indented_in"] = False
        if state["put_newline"]:
            out.append("\n")
        else:
            state["put_newline"] = True
        out.append(indent * state["depth"])
        if delta > 0:
            state["depth"] += 1
            state["indented_in"] = True

    def emit(e: Element) -> None:
        name, attrs, children, text = e
        write_indent(1)
        tag = "<" + name
        for a, v in attrs:
            if v == "" and a not in keep:
                continue
            tag += f' {a}="{_escape(v)}"'
        out.append(tag + ">")
        if text:
            out.append(_escape(text))
        for c in children:
            emit(c)
        write_indent(-1)
        out.append(f"</{name}>")

    emit(el)
    return "".join(out)


def topology(name="", type_="", dimensions="", reference="") -> Element:
    """XDMFTopology (xdmf.go:9-16): attributes name, TopologyType, Dimensions, Reference."""
    return ("Topology", [("name", name), ("TopologyType", type_), ("Dimensions", dimensions),
                         ("Reference", reference)], [], "")


def data_item(dim=0, format_="", value="", number_type="", dimensions="", precision=0, data_type="",
              endian="") -> Element:
    """XDMFDataItem (xdmf.go:27-38): attribute order Dimension, Format, NumberType, Dimensions,
    Precision, DataType, Endian; Value is character data; ints are omitted when zero."""
    return ("DataItem", [("Dimension", str(dim) if dim else ""), ("Format", format_), ("NumberType", number_type),
                         ("Dimensions", dimensions), ("Precision", str(precision) if precision else ""),
                         ("DataType", data_type), ("Endian", endian)], [], value)


def geometry(name="", items=(), type_="", reference="") -> Element:
    """XDMFGeometry (xdmf.go:18-25)."""
    return ("Geometry", [("name", name), ("Type", type_), ("Reference", reference)], list(items), "")


def _gof(v: float) -> str:
    return "%f" % v


def create_xdmf(field_names: Sequence[str], prefix: str, num: int, domain_size: Sequence[int]) -> Element:
    """CreateXDMF (xdmf.go:83-160)."""
    dim = len(domain_size)
    if dim == 2:
        dims, geo_type = f"{domain_size[0]} {domain_size[1]}", "ORIGIN_DXDY"
    elif dim == 3:
        dims, geo_type = f"{domain_size[0]} {domain_size[1]} {domain_size[2]}", "ORIGIN_DXDYDZ"
    else:
        raise ValueError("Length of domain size to be either 2 or 3")
    topo = topology(name="topo", type_=f"{dim}DCoRectMesh", dimensions=dims)
    geo = geometry(name="geo", type_=geo_type, items=[data_item(format_="XML", dimensions=str(dim), value="0.0 0.0 0.0"),
                                                      data_item(format_="XML", dimensions=str(dim), value="1.0 1.0 1.0")])
    time_str = "".join(_gof(i / num) + " " for i in range(num)) + str(num)
    grids = []
    for i in range(num):
        attrs = [("Attribute", [("Name", f), ("Center", "Node")],
                  [data_item(format_="Binary", data_type="Float", endian="Big", precision=8, dimensions=dims,
                             value=f"{prefix}_{f}_{i}.bin")], "") for f in field_names]
        grids.append(("Grid", [("Name", f"T{i + 1}"), ("GridType", "Uniform"), ("CollectionType", "")],
                      attrs + [topology(reference="/Xdmf/Domain/Topology[1]"),
                               geometry(reference="/Xdmf/Domain/Geometry[1]")], ""))
    # XDMFGrid field order: Grids, Attributes, Topology, Geometry, Time (a nil *XDMFTime is
    # skipped; struct-valued Topology / Geometry are written even when zero)
    time = ("Time", [("TimeType", "HyperSlab")],
            [data_item(format_="XML", number_type="Float", dimensions=str(dim), value=time_str)], "")
    series = ("Grid", [("Name", "TimeSeries"), ("GridType", "Collection"), ("CollectionType", "Temporal")],
              grids + [topology(), geometry(), time], "")
    return ("Xdmf", [], [("Domain", [], [topo, geo, series], "")], "")


def float64_bytes(data) -> bytes:
    return np.asarray(data, dtype=">f8").tobytes()


def real_part_as_uint8(data) -> np.ndarray:
    x = np.real(np.asarray(data)).astype(np.float64)
    mn, mx = x[0], x[0]
    for v in x:
        if v < mn:
            mn = v
        if v > mx:
            mx = v
    if abs(mx - mn) < 1e-10:
        mx = mn + 1.0
    t = 255.0 * (x - mn) / (mx - mn)
    # amd64 converts through int64: NaN becomes 0x8000000000000000, whose low byte is 0
    q = np.where(np.isnan(t), 0.0, np.trunc(t))
    return (q.astype(np.int64) & 0xFF).astype(np.uint8)
