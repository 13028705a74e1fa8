"""C-ABI library checks that need no GPU: it loads, exports every symbol the
header declares, its host-side STL readers and grid sizing match the
reference's rules, and calls without a device fail with an error code."""
import ctypes as C
import os
import re
import struct

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "lbm_capi.h")


@pytest.fixture(scope="module")
def P():
    import pv2013_amd
    pv2013_amd.lib()
    return pv2013_amd


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lbm_[a-z0-9_]+)\s*\(", text)))


def test_exports_every_declared_symbol(P):
    names = declared()
    assert len(names) >= 20
    L = P.lib()
    for n in names:
        assert hasattr(L, n), n
    from pv2013_amd import capi
    assert set(capi.EXPORTS) <= set(names)


def write_ascii_stl(path, tri):
    with open(path, "w") as fh:
        fh.write("solid t\n")
        for t in tri:
            fh.write("facet normal %.17g %.17g %.17g\n outer loop\n" % tuple(t[9:12]))
            for v in range(3):
                fh.write("  vertex %.17g %.17g %.17g\n" % tuple(t[3 * v:3 * v + 3]))
            fh.write(" endloop\nendfacet\n")
        fh.write("endsolid t\n")


def test_stl_ascii_roundtrip_and_grid(P, tmp_path):
    z = np.load(os.path.join(ROOT, "tests", "golden", "pv2013.npz"))
    tri = z["tri"]
    path = str(tmp_path / "pv.stl")
    write_ascii_stl(path, tri)
    got = P.read_stl(path)
    assert np.array_equal(got, tri)
    g = P.grid_from_bbox(got, 0.1)
    assert list(g.dims) == [421, 351, 341]
    assert g.origin[0] == tri[:, [0, 3, 6]].min() and g.raybox[5] == tri[:, [2, 5, 8]].max()


def test_stl_binary_renormalises_in_float(P, tmp_path):
    rng = np.random.default_rng(0)
    v = rng.standard_normal((5, 9)).astype(np.float32)
    n = rng.standard_normal((5, 3)).astype(np.float32) * 3
    path = str(tmp_path / "b.stl")
    with open(path, "wb") as fh:
        fh.write(b"\0" * 80 + struct.pack("<I", 5))
        for i in range(5):
            fh.write(n[i].tobytes() + v[i].tobytes() + b"\0\0")
    got = P.read_stl(path)
    norm = np.sqrt(n[:, 0] * n[:, 0] + n[:, 1] * n[:, 1] + n[:, 2] * n[:, 2]).astype(np.float32)
    assert np.array_equal(got[:, :9], v.astype(np.float64))
    assert np.array_equal(got[:, 9:], (n / norm[:, None]).astype(np.float32).astype(np.float64))


def test_generated_stl_writers_round_trip(P, tmp_path):
    """The generators' writers: an exact ASCII file reads back to the same
    doubles; a binary file (GeometryReader.cpp:119-176 layout) reads back to the
    float32 vertices, with the float32 normals renormalised in float as the
    reference's reader does."""
    from pv2013_amd.geometry import bent_pipe, write_stl_ascii, write_stl_binary
    tri, _ = bent_pipe(30, 5, 200.0)
    a, b = str(tmp_path / "a.stl"), str(tmp_path / "b.stl")
    write_stl_ascii(a, tri, exact=True)
    write_stl_binary(b, tri)
    assert np.array_equal(P.read_stl(a), tri)
    got = P.read_stl(b)
    assert got.shape == tri.shape and os.path.getsize(b) == 84 + 50 * len(tri)
    assert np.array_equal(got[:, :9], tri[:, :9].astype(np.float32).astype(np.float64))
    n = tri[:, 9:].astype(np.float32)
    norm = np.sqrt(n[:, 0] * n[:, 0] + n[:, 1] * n[:, 1] + n[:, 2] * n[:, 2]).astype(np.float32)
    assert np.array_equal(got[:, 9:], (n / norm[:, None]).astype(np.float32).astype(np.float64))


def test_header_compiles_as_c99(tmp_path):
    """include/lbm_capi.h is plain C: a C99 translation unit that takes the
    address of every declared function compiles with -pedantic -Werror (what a
    cgo or other C-FFI binding sees)."""
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if not cc:
        pytest.skip("no C compiler")
    names = declared()
    src = tmp_path / "use_capi.c"
    src.write_text('#include "lbm_capi.h"\ntypedef void (*fn_t)(void);\nfn_t all[] = {\n' +
                   "".join("    (fn_t)%s,\n" % n for n in names) + "};\nint main(void) { return all[0] == 0; }\n")
    subprocess.check_call([cc, "-std=c99", "-Wall", "-Wextra", "-pedantic", "-Werror", "-I",
                           os.path.join(ROOT, "include"), "-c", str(src), "-o", str(tmp_path / "use_capi.o")])


REF = os.path.join(ROOT, "oracle", "_ref", "ref_driver")


def ref_stl(kind, path, out):
    import subprocess
    subprocess.check_call([REF, "stl", kind, path, out])
    raw = open(out, "rb").read()
    n = int(np.frombuffer(raw[:8], np.int64)[0])
    return np.frombuffer(raw[8:], np.float64).reshape(n, 12)


@pytest.mark.skipif(not os.path.exists(REF), reason="reference build not present (needs /root/reference)")
def test_stl_readers_match_reference_build(P, oracle, tmp_path):
    """The library's and the oracle's STL readers against the reference's own
    Reader::readSTLASCII / readSTLBinary (GeometryReader.cpp:33-176), built from
    its sources: the PV2013 facets written back as ASCII, an ASCII file in
    awkward number formats, and a binary file whose normals are not unit
    length (the reference renormalises them in float)."""
    bits = lambda a: np.ascontiguousarray(a, np.float64).view(np.uint64)
    z = np.load(os.path.join(ROOT, "tests", "golden", "pv2013.npz"))
    a1 = str(tmp_path / "pv.stl")
    write_ascii_stl(a1, z["tri"])
    a2 = str(tmp_path / "odd.stl")
    with open(a2, "w") as fh:
        fh.write("solid odd\n")
        for nrm, verts in [("0 0 1", ["1e-3 +2.5 -0", "1E+02 3 4.", "0.1 0.2 0.30000000000000004"]),
                           ("-0.0 1.0e0 0", ["123456789.123456789 -1e-300 7", ".5 -.25 1e2", "3 2 1"]),
                           ("0.6 0.8 0", ["2.2250738585072014e-308 1 1", "9007199254740993 0 0", "1 2 3"])]:
            fh.write("  facet normal %s\n    outer loop\n" % nrm)
            for v in verts:
                fh.write("      vertex %s\n" % v)
            fh.write("    endloop\n  endfacet\n")
        fh.write("endsolid odd\n")
    rng = np.random.default_rng(3)
    b = str(tmp_path / "b.stl")
    with open(b, "wb") as fh:
        fh.write(b"\0" * 80 + struct.pack("<I", 200))
        for i in range(200):
            n = (rng.standard_normal(3) * 10 ** rng.uniform(-3, 3)).astype(np.float32)
            v = (rng.standard_normal(9) * 100).astype(np.float32)
            fh.write(n.tobytes() + v.tobytes() + b"\0\0")
    for kind, path in (("ascii", a1), ("ascii", a2), ("binary", b)):
        ref = ref_stl(kind, path, str(tmp_path / "ref.bin"))
        assert len(ref) > 0
        assert np.array_equal(bits(P.read_stl(path)), bits(ref)), path
        if kind == "ascii":
            assert np.array_equal(bits(oracle.read_stl_ascii(path)), bits(ref)), path


def test_calls_fail_cleanly_without_context(P):
    L = P.lib()
    h = C.c_void_p()
    rc = L.lbm_ctx_create(0, C.byref(h))
    if rc == 0:
        L.lbm_ctx_destroy(h)
        pytest.skip("a GPU is present")
    assert rc == -19
    assert L.lbm_domain_create(None, None, C.byref(h)) == -22
    assert L.lbm_step(None, 1) == -22
    assert L.lbm_get_f(None, None) == -22
    assert L.lbm_last_error(None) == b"null context"


def test_long_double_split_matches_oracle(P, oracle):
    for text in ("-175.275", "-108.379", "0.1", "-37.0", "12345.678901"):
        assert P.capi.split_long_double(text) == oracle.split_long_double(text)
