"""Writes the golden fixtures under tests/golden/ (run in a container that
has /root/reference; the fixtures are committed, this script is how they
were made).

katsura.json   Katsura-3/-4 solution vectors held by the reference's own
               test (test/katsura.f90:22-137; computed there with
               sympy/mpmath by reference/ref_katsura.py). Parsed from the
               Fortran literals as data.
c1_endpoints.npz, refsys_endpoints.npz
               Per-path endpoint records of the compiled reference
               (oracle/_ref, built by oracle/Makefile from the reference's
               own sources) for the C1 workload and for the systems of the
               reference's tests, at the default options (seed 1).
"""
import json
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
REF = os.environ.get("PHC_REFERENCE", "/root/reference")


def katsura_vectors():
    text = open(os.path.join(REF, "test", "katsura.f90")).read()
    out = {}
    for name, n, m in (("sols3", 4, 8), ("sols4", 5, 16)):
        start = text.index(f"{name} = reshape([")
        end = text.index("])", start)
        body = text[start:end]
        pairs = re.findall(r"\(\s*([-+0-9.eE]+)_wp\s*,\s*([-+0-9.eE]+)_wp\s*\)", body)
        vals = [[float(a), float(b)] for a, b in pairs]
        assert len(vals) == n * m, (name, len(vals))
        out[name] = vals  # column-major n x m, (re, im)
    return out


def main():
    import oracle
    from polyhomocont_amd import abi, workloads
    with open(os.path.join(HERE, "katsura.json"), "w") as f:
        json.dump(katsura_vectors(), f)
    R = oracle.ref()
    opts = abi.default_options()
    rec = {}
    systems = {"c1": workloads.config_system("c1")}
    systems.update(workloads.ref_test_systems())
    del systems["cyclic6"]
    for name, usys in systems.items():
        e = R.endpoints(usys, opts)
        g, b, a, ns = R.homotopy(usys, opts.seed)
        rec[name + "/status"] = e.status[:ns]
        rec[name + "/z"] = e.z[:ns]
        rec[name + "/winding"] = e.winding[:ns]
        rec[name + "/eg_status"] = e.eg_status[:ns]
        rec[name + "/eg_error"] = e.eg_error[:ns]
        rec[name + "/naccepted"] = e.naccepted[:ns]
        rec[name + "/nrejected"] = e.nrejected[:ns]
        rec[name + "/t"] = e.t[:ns]
        rec[name + "/gamma"] = np.array([g])
        rec[name + "/b"] = b
        rec[name + "/patch"] = a
    np.savez_compressed(os.path.join(HERE, "endpoints.npz"), **rec)


if __name__ == "__main__":
    main()
