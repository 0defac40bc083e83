"""The C ABI: the library loads, exports every declared symbol, and the
ctypes mirror of ``nz_problem`` has the C layout.  No compute calls (CPU)."""

from __future__ import annotations

import ctypes as C
import os
import re
import subprocess

import pytest

from conftest import ROOT, nz_module

HEADER = os.path.join(ROOT, "include", "nozzle_hip.h")


def declared_functions() -> list[str]:
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return re.findall(r"\b(nz_[a-z_]+)\s*\(", text)


def test_library_exports_every_declared_symbol():
    lib = nz_module("_lib")
    if not os.path.exists(lib.LIB_PATH):
        pytest.fail("libnozzle_hip.so is not built; run __graft_entry__.build()")
    dll = C.CDLL(lib.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 9
    for n in names:
        assert hasattr(dll, n), n
        assert n in lib.SIGNATURES, n


@pytest.mark.parametrize("cname,pyname", [("nz_problem", "Problem"), ("nz_mesh_desc", "MeshDesc"),
                                           ("nz_edge_maps", "EdgeMaps"),
                                           ("nz_geometry_desc", "GeometryDesc"),
                                           ("nz_geometry_out", "GeometryOut"),
                                           ("nz_quasi1d_desc", "Quasi1DDesc")])
def test_struct_layout(tmp_path, cname, pyname):
    """Compile a probe against the header and compare offsets with ctypes."""
    lib = nz_module("_lib")
    mirror = getattr(lib, pyname)
    fields = [f for f, _ in mirror._fields_]
    src = tmp_path / "probe.c"
    body = "\n".join(f'printf("%s %zu\\n", "{f}", offsetof({cname}, {f}));' for f in fields)
    src.write_text(
        "#include <stdio.h>\n#include <stddef.h>\n#include \"nozzle_hip.h\"\n"
        "int main(void){" + body + f'printf("size %zu\\n", sizeof({cname})); return 0;}}\n'
    )
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    got = dict(line.split() for line in out if line)
    for f in fields:
        assert int(got[f]) == getattr(mirror, f).offset, f
    assert int(got["size"]) == C.sizeof(mirror)


def test_unsupported_configuration_is_an_error():
    """An unknown element configuration is rejected on the host, before any launch."""
    lib = nz_module("_lib")
    if not os.path.exists(lib.LIB_PATH):
        pytest.fail("libnozzle_hip.so is not built")
    dll = lib.load()
    P = lib.Problem()
    P.n_elem, P.nface, P.nbf, P.nqv, P.nqf = 10, 3, 7, 5, 3  # no such element
    rc = dll.nz_residual(C.byref(P), None, None, None, None)
    assert rc == -1
    assert b"unsupported element configuration" in dll.nz_last_error()
    assert dll.nz_scratch_doubles(C.byref(P)) == -1


@pytest.mark.parametrize("entry", ["nz_residual", "nz_rate", "nz_rate_noalloc"])
@pytest.mark.parametrize("chain", [1, 2, 3])
def test_stage_to_stage_boundary_fluxes_are_refused_outside_stages(entry, chain):
    """``bflux_chain`` belongs to ``nz_stage``: the residual and rate entry
    points refuse it on the host, before any launch (a rate's epilogue would
    otherwise write boundary fluxes of the rate into the workspace)."""
    lib = nz_module("_lib")
    if not os.path.exists(lib.LIB_PATH):
        pytest.fail("libnozzle_hip.so is not built")
    dll = lib.load()
    P = lib.Problem()
    P.n_elem, P.nface, P.nbf, P.nqv, P.nqf, P.affine, P.straight = 10, 3, 3, 4, 2, 1, 1
    P.bflux_chain = chain
    rc = getattr(dll, entry)(C.byref(P), None, None, None, None)
    assert rc == -1
    assert b"bflux_chain is for nz_stage only" in dll.nz_last_error()
    P.bflux_chain = 0  # the same problem passes the check and stops at the null arrays
    assert getattr(dll, entry)(C.byref(P), None, None, None, None) == -1
    assert b"null array" in dll.nz_last_error()


def test_face_nodes_match_the_reference_tables():
    """nz_face_node (a host query) against the reference's face tables: for
    every element kind, order and geometry order, and both sides, the rows of
    phi_face that are not round-off are exactly the p + 1 listed nodes, in
    lattice order.  Where the kernels drop the other rows (triangles, and
    quadrilaterals up to p = 2) those rows are at most 2.6e-15; p = 3
    quadrilaterals, which keep the full sums, carry up to 4.3e-14 there."""
    import numpy as np

    lib = nz_module("_lib")
    dll = C.CDLL(lib.LIB_PATH)
    fn = dll.nz_face_node
    fn.restype, fn.argtypes = C.c_int, [C.c_int] * 4
    tabs = np.load(os.path.join(ROOT, "tests", "golden", "tables.npz"))
    seen = 0
    for kind, nface in (("tri", 3), ("quad", 4)):
        for p in (1, 2, 3):
            for q in (1, 2):
                pf = tabs[f"{kind}_p{p}_q{q}_phi_face"]
                assert pf.shape[1] == nface
                for f in range(nface):
                    nodes = [fn(nface, p, f, j) for j in range(p + 1)]
                    off = np.ones(pf.shape[2], bool)
                    off[nodes] = False
                    for side in range(2):
                        big = np.flatnonzero(np.abs(pf[side, f]).max(axis=1) > 1e-13)
                        assert sorted(nodes) == list(big), (kind, p, q, f, side, nodes, big)
                        small = np.abs(pf[side, f][off]).max()
                        bound = 4.3e-14 if (kind == "quad" and p == 3) else 2.6e-15
                        assert small <= bound, (kind, p, q, f, side, small)
                    seen += 1
    assert seen == 2 * 3 * (3 + 4)
    assert fn(3, 0, 0, 0) == -1 and fn(3, 2, 3, 0) == -1 and fn(4, 2, 0, 3) == -1
