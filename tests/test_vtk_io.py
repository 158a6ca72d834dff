"""VTK XML meshes and faces (.vtu / .vtp, the formats svFSI's vtkXMLParser
reads) through the C ABI, on the host: files written here in every
encoding the reader accepts (ascii; inline base64; appended raw and
base64; UInt32 and UInt64 headers; zlib-compressed blocks; Float32 and
Float64 points; Int32 and Int64 connectivity) read back exactly, GlobalNodeID
and GlobalElementID included.  The writer follows the VTK XML file-format
document: header = byte count, or for vtkZLibDataCompressor
[nblocks, blocksize, lastblocksize, compressed sizes...] then the blocks."""
import base64
import struct
import zlib

import numpy as np
import pytest

from svfsi_amd import mesh, meshio
from svfsi_amd._lib import SvfsiError

VT = {np.dtype("float64"): "Float64", np.dtype("float32"): "Float32", np.dtype("int32"): "Int32",
      np.dtype("int64"): "Int64", np.dtype("uint8"): "UInt8"}


def payload(a, hdr, comp, block=1000):
    raw = np.ascontiguousarray(a).tobytes()
    h = "<I" if hdr == "UInt32" else "<Q"
    if not comp:
        return struct.pack(h, len(raw)) + raw, None
    blocks = [raw[i:i + block] for i in range(0, len(raw), block)] or [b""]
    cz = [zlib.compress(b) for b in blocks]
    last = len(blocks[-1]) if len(blocks[-1]) != block else 0
    head = struct.pack("<" + h[1] * (3 + len(cz)), len(cz), block, last, *[len(c) for c in cz])
    return head, b"".join(cz)


def write_vtk(path, x, cells, ctype, fmt, hdr="UInt32", comp=False, poly=False, gnid=None, geid=None,
              xdtype=np.float64, idtype=np.int32):
    arrays = {"pts": x.astype(xdtype), "conn": cells.astype(idtype).ravel(),
              "offs": (np.arange(1, len(cells) + 1) * cells.shape[1]).astype(idtype)}
    if not poly:
        arrays["types"] = np.full(len(cells), ctype, dtype=np.uint8)
    if gnid is not None:
        arrays["gn"] = gnid.astype(np.int32)
    if geid is not None:
        arrays["ge"] = geid.astype(np.int32)
    appended = b""

    def da(key, name, ncomp=1):
        nonlocal appended
        a = arrays[key]
        attrs = f'type="{VT[a.dtype]}" Name="{name}" NumberOfComponents="{ncomp}"'
        if fmt == "ascii":
            return f'<DataArray {attrs} format="ascii">\n{" ".join(map(repr, a.ravel().tolist()))}\n</DataArray>'
        head, body = payload(a, hdr, comp)
        if fmt == "binary":
            txt = base64.b64encode(head).decode() + (base64.b64encode(body).decode() if body is not None else "")
            return f'<DataArray {attrs} format="binary">\n{txt}\n</DataArray>'
        off = len(appended)
        blob = head + (body or b"")
        appended += base64.b64encode(blob) if fmt == "appended64" else blob
        return f'<DataArray {attrs} format="appended" offset="{off}"/>'

    kind = "PolyData" if poly else "UnstructuredGrid"
    sec = "Polys" if poly else "Cells"
    cellarrays = da("conn", "connectivity") + da("offs", "offsets") + ("" if poly else da("types", "types"))
    pd = f"<PointData>{da('gn', 'GlobalNodeID')}</PointData>" if gnid is not None else ""
    cd = f"<CellData>{da('ge', 'GlobalElementID')}</CellData>" if geid is not None else ""
    ncell = f'NumberOf{"Polys" if poly else "Cells"}="{len(cells)}"'
    xml = (f'<?xml version="1.0"?>\n<VTKFile type="{kind}" version="1.0" byte_order="LittleEndian" '
           f'header_type="{hdr}"' + (' compressor="vtkZLibDataCompressor"' if comp else "") + ">\n"
           f'<{kind}>\n<Piece NumberOfPoints="{len(x)}" {ncell}>\n{pd}{cd}'
           f"<Points>{da('pts', 'Points', 3)}</Points>\n<{sec}>{cellarrays}</{sec}>\n</Piece>\n</{kind}>\n")
    data = xml.encode()
    if fmt.startswith("appended"):
        enc = "base64" if fmt == "appended64" else "raw"
        data += f'<AppendedData encoding="{enc}">\n_'.encode() + appended + b"\n</AppendedData>\n"
    data += b"</VTKFile>\n"
    with open(path, "wb") as f:
        f.write(data)


@pytest.mark.parametrize("fmt", ["ascii", "binary", "appended", "appended64"])
@pytest.mark.parametrize("hdr,comp", [("UInt32", False), ("UInt64", False), ("UInt32", True), ("UInt64", True)])
def test_vtu_round_trip(tmp_path, fmt, hdr, comp):
    if fmt == "ascii" and (hdr != "UInt32" or comp):
        pytest.skip("headers and compression do not apply to ascii arrays")
    m = mesh.cylinder_tet4(0.6)
    gn = np.arange(1, m.nNo + 1) * 3
    ge = np.arange(1, m.nEl + 1)
    p = str(tmp_path / "m.vtu")
    write_vtk(p, m.x, m.ien, 10, fmt, hdr, comp, gnid=gn, geid=ge)
    x, cells, gnr, ger = meshio.read_vtk(p)
    assert np.array_equal(x, m.x) and np.array_equal(cells, m.ien)
    assert np.array_equal(gnr, gn) and np.array_equal(ger, ge)


def test_vtu_float32_int64_tet10_and_vtp_face(tmp_path):
    m = mesh.cube_tet10(2)
    p = str(tmp_path / "m10.vtu")
    write_vtk(p, m.x, m.ien, 24, "binary", "UInt64", True, xdtype=np.float32, idtype=np.int64)
    x, cells, gn, ge = meshio.read_vtk(p)
    assert np.array_equal(x, m.x.astype(np.float32).astype(np.float64)) and np.array_equal(cells, m.ien)
    assert gn is None and ge is None
    f, e = mesh.boundary_faces(m.ien[:, :4])
    nodes, local = np.unique(f, return_inverse=True)
    q = str(tmp_path / "face.vtp")
    write_vtk(q, m.x[nodes], local.reshape(f.shape), 5, "appended", "UInt32", True, poly=True,
              gnid=nodes + 1, geid=e + 1)
    xf, cf, gnf, gef = meshio.read_vtk(q)
    assert np.array_equal(gnf[cf], f + 1) and np.array_equal(gef, e + 1)
    assert np.array_equal(xf, m.x[nodes])


BAD_HEADERS = {
    # uncompressed: the byte count of the array
    False: [lambda h, n: struct.pack(h, n), ],
    # zlib: [nblocks, blocksize, lastblocksize, compressed sizes...]
    True: [lambda h, n: struct.pack("<" + h[1] * 4, n, 1000, 0, 10),     # block count
           lambda h, n: struct.pack("<" + h[1] * 4, 1, n, 0, 10),        # block size
           lambda h, n: struct.pack("<" + h[1] * 4, 1, 1000, 8, n),      # compressed size
           lambda h, n: struct.pack("<" + h[1] * 4, 2, 1000, 0, 10)],    # sizes beyond the data
}


@pytest.mark.parametrize("fmt", ["binary", "appended", "appended64"])
@pytest.mark.parametrize("hdr", ["UInt32", "UInt64"])
@pytest.mark.parametrize("comp", [False, True])
def test_corrupt_block_headers_are_errors(tmp_path, monkeypatch, fmt, hdr, comp):
    """Sizes read from a block header that exceed the file (including values
    near 2^32 and 2^64, where an unchecked offset sum wraps) give an error
    return, never an abort of the calling process."""
    m = mesh.cube_tet4(1)
    h = "<I" if hdr == "UInt32" else "<Q"
    huge = (1 << 32) - 1 if hdr == "UInt32" else (1 << 64) - 1
    for make in BAD_HEADERS[comp]:
        for n in (huge, huge - 7, 1 << 31):
            def bad(a, hdr_, comp_, block=1000, make=make, n=n):
                return make(h, n), (b"" if comp_ else None)
            monkeypatch.setattr(__import__(__name__), "payload", bad, raising=True)
            p = str(tmp_path / f"bad_{fmt}_{hdr}_{comp}.vtu")
            write_vtk(p, m.x, m.ien, 10, fmt, hdr, comp)
            with pytest.raises(SvfsiError):
                meshio.read_vtk(p)


@pytest.mark.parametrize("fmt", ["binary", "appended", "appended64"])
@pytest.mark.parametrize("comp", [False, True])
def test_truncated_files_are_errors(tmp_path, fmt, comp):
    """Every prefix cut of a valid file is either read or refused with an
    error; none crashes the reader."""
    m = mesh.cube_tet4(1)
    p = tmp_path / "m.vtu"
    write_vtk(str(p), m.x, m.ien, 10, fmt, "UInt64", comp)
    data = p.read_bytes()
    q = tmp_path / "cut.vtu"
    for cut in range(0, len(data), max(1, len(data) // 97)):
        q.write_bytes(data[:cut])
        try:
            meshio.read_vtk(str(q))
        except SvfsiError:
            pass


def test_appended_offset_beyond_the_file_is_an_error(tmp_path):
    m = mesh.cube_tet4(1)
    for fmt in ("appended", "appended64"):
        p = tmp_path / f"m_{fmt}.vtu"
        write_vtk(str(p), m.x, m.ien, 10, fmt, "UInt32", False)
        s = p.read_bytes().replace(b'offset="0"', b'offset="99999999999"', 1)
        p.write_bytes(s)
        with pytest.raises(SvfsiError):
            meshio.read_vtk(str(p))
