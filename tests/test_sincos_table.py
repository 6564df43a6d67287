"""polyhomocont_amd/csrc/phc_sincos_table.h holds sin(i/128), cos(i/128) as
double-double pairs. The header must equal what its generator writes, and
each pair must be the nearest double plus the nearest double to the
remainder (checked here with the decimal module at 60 digits)."""
import importlib.util
import os
import re
from decimal import Decimal, getcontext

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, "polyhomocont_amd", "csrc", "phc_sincos_table.h")


def _table():
    text = open(HDR).read()
    body = text[text.index("{") + 1:text.index("};")]
    vals = [float.fromhex(t) for t in re.findall(r"-?0x[0-9a-fA-F.]+p[+-]?\d+", body)]
    assert len(vals) == 512
    return vals


def _sin_cos(x):
    """Taylor series of sin and cos at 60 significant digits."""
    getcontext().prec = 60
    s, c = Decimal(0), Decimal(0)
    term = Decimal(1)
    for k in range(60):
        if k > 0:
            term = term * x / k
        r = k % 4
        if r == 0:
            c += term
        elif r == 1:
            s += term
        elif r == 2:
            c -= term
        else:
            s -= term
    return s, c


def _split(v):
    hi = float(v)
    return hi, float(v - Decimal(hi))


def test_table_entries_are_rounded_double_doubles():
    vals = _table()
    for i in range(128):
        s, c = _sin_cos(Decimal(i) / 128)
        assert (vals[4 * i], vals[4 * i + 1]) == _split(s), i
        assert (vals[4 * i + 2], vals[4 * i + 3]) == _split(c), i


def test_generator_reproduces_header(tmp_path, monkeypatch):
    if importlib.util.find_spec("mpmath") is None:
        pytest.skip("mpmath is not installed")
    gen = os.path.join(ROOT, "polyhomocont_amd", "csrc", "gen_sincos_table.py")
    spec = importlib.util.spec_from_file_location("gen_sincos_table", gen)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    monkeypatch.setattr(mod, "HERE", str(tmp_path))
    mod.main()
    assert open(tmp_path / "phc_sincos_table.h").read() == open(HDR).read()
