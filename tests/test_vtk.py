"""VTK ImageData output (ps_write_vti), the writer of the reference's test
programs (c++/test/unbounded3d/sources/main.cpp:329-455).  Host code: runs
without a GPU."""
import xml.etree.ElementTree as ET

import numpy as np

import poissonsolver_amd as P


def _fields(n):
    rng = np.random.default_rng(2)
    a = [rng.standard_normal(n[::-1]) for _ in range(3)]
    b = rng.standard_normal(n[::-1])
    return {"B": (b, None, None), "A": tuple(a)}


def _expect(f):
    return np.stack([np.zeros(f[0].size) if c is None else c.reshape(-1) for c in f], 1)


def test_ascii_piece_and_pvti(tmp_path):
    n = (5, 4, 3)
    f = _fields(n)
    P.write_vti(tmp_path, "mesh", n, (-1.0, -1.0, -1.0), (0.4, 0.5, 2 / 3), f)
    root = ET.parse(tmp_path / "mesh_P00.vti").getroot()
    img = root.find("ImageData")
    assert img.get("WholeExtent").split() == ["0", "5", "0", "4", "0", "3"]
    assert [float(v) for v in img.get("Spacing").split()] == [0.4, 0.5, float("%.8g" % (2 / 3))]
    arrs = img.find("Piece").find("CellData").findall("DataArray")
    assert [a.get("Name") for a in arrs] == ["B", "A"]
    for a in arrs:
        vals = np.array(a.text.split(), dtype=float).reshape(-1, 3)
        want = _expect(f[a.get("Name")])
        assert np.allclose(vals, want, rtol=1e-8, atol=0)     # %17.8e, as the reference writes
    p = ET.parse(tmp_path / "mesh.pvti").getroot()
    pimg = p.find("PImageData")
    assert [x.get("Name") for x in pimg.find("PCellData")] == ["B", "A"]
    assert pimg.find("Piece").get("Source") == "./mesh_P00.vti"


def test_binary_appended_roundtrip(tmp_path):
    n = (7, 3, 2)
    f = _fields(n)
    P.write_vti(tmp_path, "out", n, (0, 0, 0), (1, 1, 1), f, binary=True)
    raw = (tmp_path / "out_P00.vti").read_bytes()
    head, _, rest = raw.partition(b"<AppendedData encoding='raw'>\n   _")
    root = ET.fromstring(head + b"</VTKFile>")
    arrs = root.find("ImageData").find("Piece").find("CellData").findall("DataArray")
    for a in arrs:
        off = int(a.get("offset"))
        nbytes = int(np.frombuffer(rest[off:off + 8], np.uint64)[0])
        vals = np.frombuffer(rest[off + 8:off + 8 + nbytes], np.float64).reshape(-1, 3)
        assert np.array_equal(vals, _expect(f[a.get("Name")]))
