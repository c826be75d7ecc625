"""VTU output (host only): the bytes equal an independent restatement of the
reference writer (data_writer.cpp:93-244) on the same mesh and arrays, and the
file reads back to the arrays it was given."""
import re
import struct

import numpy as np
import pytest

NAMES = ["RHO", "RHOU_X", "RHOU_Y", "RHOE", "U_X", "U_Y", "P", "T", "H", "CFL"]


def reference_bytes(n_nodes, nodes_of_cell, node_coords, n_cells, arrays):
    """data_writer.cpp:93-244 restated: header with running offsets, then the
    raw appended block (4-byte length + data per array)."""
    h = ['<?xml version="1.0"?>\n', '<VTKFile type="UnstructuredGrid" version="0.1" byte_order="LittleEndian">\n',
         "  <UnstructuredGrid>\n", '    <Piece NumberOfPoints="%d" NumberOfCells="%d">\n' % (n_nodes, n_cells),
         "      <PointData>\n", "      </PointData>\n", "      <CellData>\n"]
    off = 0
    for name, _ in arrays:
        h.append('        <DataArray type="Float64" Name="%s" format="appended" offset="%d">\n' % (name, off))
        h.append("        </DataArray>\n")
        off += 4 + n_cells * 8
    h += ["      </CellData>\n", "      <Points>\n",
          '        <DataArray type="Float64" NumberOfComponents="3" format="appended" offset="%d">\n' % off,
          "        </DataArray>\n", "      </Points>\n", "      <Cells>\n"]
    off += 4 + n_nodes * 3 * 8
    h += ['        <DataArray type="UInt32" Name="connectivity" format="appended" offset="%d">\n' % off,
          "        </DataArray>\n"]
    off += 4 + len(nodes_of_cell) * 4
    h += ['        <DataArray type="UInt32" Name="offsets" format="appended" offset="%d">\n' % off,
          "        </DataArray>\n"]
    off += 4 + n_cells * 4
    h += ['        <DataArray type="UInt8" Name="types" format="appended" offset="%d">\n' % off,
          "        </DataArray>\n", "      </Cells>\n", "    </Piece>\n", "  </UnstructuredGrid>\n",
          '<AppendedData encoding="raw">\n_']
    b = "".join(h).encode()
    for _, a in arrays:
        b += struct.pack("<I", n_cells * 8) + np.asarray(a, "<f8").tobytes()
    pts = np.zeros((n_nodes, 3))
    pts[:, :2] = node_coords.reshape(-1, 2)
    b += struct.pack("<I", n_nodes * 24) + pts.astype("<f8").tobytes()
    b += struct.pack("<I", len(nodes_of_cell) * 4) + np.asarray(nodes_of_cell, "<u4").tobytes()
    b += struct.pack("<I", n_cells * 4) + (np.arange(1, n_cells + 1, dtype="<u4") * 3).tobytes()
    b += struct.pack("<I", n_cells) + bytes([7]) * n_cells
    return b + b"\n</AppendedData>\n</VTKFile>\n"


@pytest.mark.parametrize("variables", [NAMES, ["P", "RHO", "CFL"]])
def test_vtu_bytes_match_reference_writer(mlr, tmp_path, variables):
    L = mlr.Layout(mlr.mesh_gen("cartesian_tri", 7, 5, 1.0, 1.0, 0.2, 3), mlr.numerics(recon="FO"))
    a = L.mesh_arrays()
    N = L.n_cells
    rng = np.random.default_rng(0)
    U, W, cfl = rng.normal(size=(N, 4)), rng.normal(size=(N, 5)), rng.random(N)
    path = tmp_path / "out_000010.vtu"
    L.write_vtu(path, variables, U, W, cfl)
    cols = dict(zip(NAMES, [U[:, i] for i in range(4)] + [W[:, i] for i in range(5)] + [cfl]))
    ref = reference_bytes(L.n_nodes, a["nodes_of_cell"], a["node_coords"], N, [(v, cols[v]) for v in variables])
    got = path.read_bytes()
    assert got == ref
    # and it reads back
    head, _, body = got.partition(b'<AppendedData encoding="raw">\n_')
    offs = [int(o) for o in re.findall(rb'offset="(\d+)"', head)]
    for v, o in zip(variables, offs):
        n = struct.unpack_from("<I", body, o)[0]
        assert np.array_equal(np.frombuffer(body, "<f8", n // 8, o + 4), cols[v])


def test_write_data_table_errors(mlr, tmp_path):
    base = open(__file__.replace("test_vtu.py", "golden/examples/sod.toml")).read()
    bad = base.replace('prefix = "./solut/all/sod_all"\n', "")
    p = tmp_path / "bad.toml"
    p.write_text(bad)
    with pytest.raises(mlr.MallardError, match="DataWriter: prefix not specified."):
        mlr.parse_config(p)
    p.write_text(base.replace('"CFL", "RHO"', '"CFL", "DENSITY"'))
    with pytest.raises(mlr.MallardError, match="DataWriter: Unknown variable: DENSITY."):
        mlr.parse_config(p)
    p.write_text(base.replace('format = "vtu"', 'format = "hdf5"'))
    with pytest.raises(mlr.MallardError, match="DataWriter: Unknown format type: hdf5."):
        mlr.parse_config(p)
