"""The C-ABI library loads and exports every entry point include/gpf.h declares (CPU only)."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    syms = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(gpf_[a-z0-9_]+)\s*\(", text):
            syms.add(m.group(1))
    return sorted(syms)


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "gpf_solver_propagate" in syms and "gpf_fft_forward" in syms and len(syms) > 30


def test_library_exports_every_declared_symbol():
    import gopf_amd
    L = gopf_amd.lib()
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing
    # exported with C linkage (no C++ mangling)
    out = os.popen(f"nm -D --defined-only {gopf_amd.LIB_PATH}").read()
    for s in declared_symbols():
        assert re.search(rf"\bT {s}$", out, flags=re.M), s


def test_host_only_calls_work_without_gpu():
    """Model building / parsing and the Freq helpers never touch the device."""
    import gopf_amd as G
    m = G.Model()
    m.AddField("conc", n=16)
    m.AddScalar("k", 2.0)
    m.AddEquation("dconc/dt = k*LAP conc - LAP conc^3")
    m.Init()
    assert m.AllFieldNames() == ["conc", "conc^3"]
    assert G.Freq([8, 16], 17) == [0.125, 0.0625]
    assert G.ConjugateNode([8, 8], 9) == 63
    v = G.NewVandeven(4)
    assert abs(v.Eval(0.0) - 1.0) < 1e-12


def test_errors_are_reported_not_raised_as_crashes():
    import gopf_amd as G
    m = G.Model()
    m.AddField("conc", n=4)
    with pytest.raises(G.GpfError) as ei:
        m.AddEquation("conc = LAP conc")
    assert ei.value.code == 2 and "leibniz" in str(ei.value)
    with pytest.raises(G.GpfError):
        G.Solver(m, [2, 2], 0.1)  # model not initialised


def test_fft_rejects_a_buffer_of_the_wrong_length():
    """FFTWWrapper.FFT/IFFT copy between the plan's buffer and the caller's slice
    (pfutil/fftWrap.go:26-39); the C entry points take the slice's length and refuse any
    length other than the plan's size before touching memory or the device."""
    import gopf_amd as G
    L = G.lib()
    h = ctypes.c_void_p()
    dims = (ctypes.c_int * 2)(8, 16)
    assert L.gpf_fft_create(dims, 2, G.GPF_FP64, 0, ctypes.byref(h)) == 0
    try:
        short = (ctypes.c_double * (2 * 64))()
        for fn in (L.gpf_fft_forward, L.gpf_fft_inverse):
            assert fn(h, short, 64) == G.GPF_EINVAL
            assert "does not match the plan" in L.gpf_last_error().decode()
            assert fn(h, short, 129) == G.GPF_EINVAL
        ft = G.FourierTransform([8, 16])
        with pytest.raises(G.GpfError) as ei:
            ft.FFT(np.zeros(100, dtype=np.complex128))
        assert ei.value.code == G.GPF_EINVAL
    finally:
        L.gpf_fft_destroy(h)


def test_fft_lengths_beyond_the_plans_are_refused():
    """Lengths above 1024 without a compiled schedule (here the prime 1031) have no plan;
    NewFFTW's counterpart refuses them on the host, before any device call."""
    import gopf_amd as G
    with pytest.raises(G.GpfError) as ei:
        G.FourierTransform([1031, 4])
    assert ei.value.code == G.GPF_EINVAL and "1031" in str(ei.value)
