"""Mesh files, result CSVs and restart values (host code of libbsim.so, no
GPU): the reference's mesh_file_test expectations on its own fixture files
(copied under tests/golden/meshes from the reference's tests/data), exact
write/read round trips, the result formats of parallel_writer.cpp and the
restart reader of field_reader.cpp with its error cases."""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

import oracle_lib as orc
from babelsim_amd import MeshData, read_field_values, write_field_csv, write_result_metadata

MESHES = Path(__file__).resolve().parent / "golden" / "meshes"


def geometry(mesh: MeshData):
    return orc.oracle_mesh(mesh.arrays()).geometry()


def test_reference_mesh_file_expectations():
    """tests/mesh_file_test.cpp:11-47 on the reference's own fixtures."""
    mesh = MeshData.read(MESHES / "babelsim_channel.mesh")
    nv, nf, _, npatch, nc = mesh.sizes()
    assert nc == 9 and npatch == 6
    g = geometry(mesh)
    np.testing.assert_allclose(g["V"], 1.0, rtol=1e-12)
    np.testing.assert_allclose(g["C"][0], (0.5, 0.5, 0.5), rtol=1e-12)
    kinds = [k for _, k in mesh.patches()]
    assert kinds.count("wall") == 2 and kinds.count("inlet") == 1 and kinds.count("outlet") == 1
    assert kinds.count("symmetry") == 2
    one = MeshData.read(MESHES / "babelsim_vertices.mesh")
    assert one.sizes()[4] == 1
    assert geometry(one)["V"][0] == pytest.approx(1.0, rel=1e-12)
    concave = MeshData.read(MESHES / "babelsim_concave.mesh")
    a = concave.arrays()
    assert concave.sizes()[4] == 1 and len(a["owner"]) == 8 and a["fpo"][1] - a["fpo"][0] == 6
    assert geometry(concave)["V"][0] == pytest.approx(3.0, rel=1e-12)


def test_mesh_write_read_round_trip(tmp_path):
    mesh = MeshData.hex((5, 4, 3), jitter=0.2, split=0.3, seed=11)
    path = tmp_path / "m.mesh"
    mesh.write(path)
    back = MeshData.read(path)
    a, b = mesh.arrays(), back.arrays()
    for key in ("fpo", "fpi", "owner", "neighbour", "patch"):
        np.testing.assert_array_equal(a[key], b[key])
    np.testing.assert_array_equal(a["vertices"], b["vertices"])  # 17 digits: exact
    assert back.patches()[0] == ("patch0", "generic")
    reread = tmp_path / "again.mesh"
    back.write(reread)
    assert reread.read_text() == path.read_text()


@pytest.mark.parametrize("edit,match", [
    (lambda t: t.replace("BABELSIM_MESH 3", "BABELSIM_MESH 2"), "version 3"),
    (lambda t: t.rstrip().rsplit("\n", 1)[0] + "\n", "end"),
    (lambda t: t + "extra\n", "trailing"),
    (lambda t: t.replace("symmetry", "slippery", 1), "unknown patch kind"),
    (lambda t: t.replace("face 4 0 -1 0 0 4 6 2", "face 4 7 -1 0 0 4 6 2"), "no faces"),
    (lambda t: t.replace("face 4 0 -1 0 0 4 6 2", "face 4 0 -1 0 0 4 6 4"), "repeats"),
    (lambda t: t.replace("face 4 0 -1 0 0 4 6 2", "face 4 0 -1 -1 0 4 6 2"), "patch"),
    (lambda t: t.replace("face 4 0 -1 0 0 4 6 2", "face 4 0 -1 0 0 4 6 99"), "vertex index"),
])
def test_invalid_mesh_files_are_rejected(tmp_path, edit, match):
    text = (MESHES / "babelsim_vertices.mesh").read_text()
    bad = tmp_path / "bad.mesh"
    bad.write_text(edit(text))
    with pytest.raises(RuntimeError, match=match):
        MeshData.read(bad)


def test_result_csv_format_and_restart_round_trip(tmp_path):
    """parallel_writer.cpp:42-75 format; field_reader.cpp:143-241 reads it back
    by global id from the rank directory, from the file and from id/value rows."""
    rng = np.random.default_rng(3)
    n = 50
    centres = rng.random((n, 3))
    U = rng.standard_normal((n, 3))
    ids = rng.permutation(n).astype(np.int32)  # rows in some other (device) order
    write_field_csv(tmp_path / "final", "U", U, centres, global_ids=ids)
    csv = tmp_path / "final" / "rank-0000" / "U.csv"
    lines = csv.read_text().splitlines()
    assert lines[0] == "global_id,x,y,z,value0,value1,value2" and len(lines) == n + 1
    first = lines[1].split(",")
    assert int(first[0]) == ids[0] and float(first[4]) == U[0, 0]  # 17 significant digits round-trip
    expect = np.empty_like(U)
    expect[ids] = U
    np.testing.assert_array_equal(read_field_values(tmp_path / "final", n, 3, name="U"), expect)
    np.testing.assert_array_equal(read_field_values(csv, n, 3), expect)
    p = rng.standard_normal(n)
    data = tmp_path / "p.data"
    data.write_text("# restart values\n" + "".join(f"{i} {float(p[i])!r}  # cell {i}\n" for i in range(n)))
    np.testing.assert_array_equal(read_field_values(data, n, 1), p)


@pytest.mark.parametrize("rows,match", [
    (lambda n: [f"{i} 1.0" for i in range(n - 1)], "missing"),
    (lambda n: [f"{i} 1.0" for i in range(n)] + ["0 2.0"], "duplicate"),
    (lambda n: [f"{i} 1.0" for i in range(n)] + [f"{n} 2.0"], "invalid global cell"),
    (lambda n: [f"{i} 1.0 7" for i in range(n)], "unexpected data"),
])
def test_restart_reader_errors(tmp_path, rows, match):
    data = tmp_path / "v.data"
    data.write_text("\n".join(rows(6)) + "\n")
    with pytest.raises(RuntimeError, match=match):
        read_field_values(data, 6, 1)


def test_metadata_and_mesh_geometry(tmp_path):
    """metadata.bs lines and mesh.geometry rows (id, vertex count, the sorted
    distinct vertices of the cell's faces: src/core/mesh.cpp:473-484)."""
    mesh = MeshData.read(MESHES / "babelsim_vertices.mesh")
    write_result_metadata(tmp_path / "t1", mesh, "1.5", [("U", "vector", "cell"), ("p", "scalar", "cell")])
    meta = (tmp_path / "t1" / "rank-0000" / "metadata.bs").read_text().splitlines()
    assert meta == ["format babelsim_result 3", "time 1.5", "rank 0", "ranks 1", "global_cell_count 1",
                    "owned_cells 1", "field U vector cell", "field p scalar cell"]
    geo = (tmp_path / "t1" / "rank-0000" / "mesh.geometry").read_text().strip().split(",")
    assert geo[:2] == ["0", "8"]
    verts = np.array(geo[2:], float).reshape(8, 3)
    np.testing.assert_array_equal(verts, mesh.arrays()["vertices"].reshape(8, 3))


def test_renumbered_output_keeps_original_ids(tmp_path):
    """The device numbering (reverse Cuthill-McKee) never reaches a file:
    values written with the renumbering's new-to-old map as global ids read
    back in the caller's cell order."""
    mesh = MeshData.hex((6, 5, 4), jitter=0.1, seed=2)
    prepared, new_to_old, _ = mesh.prepare(renumber=True)
    n = mesh.sizes()[4]
    assert sorted(new_to_old.tolist()) == list(range(n)) and not np.array_equal(new_to_old, np.arange(n))
    g = geometry(prepared)
    values = g["C"][:, 0] * 10 + g["C"][:, 1]  # a function of position, in the new order
    write_field_csv(tmp_path / "r", "f", values, g["C"], global_ids=new_to_old)
    back = read_field_values(tmp_path / "r", n, 1, name="f")
    c0 = geometry(mesh)["C"]
    np.testing.assert_allclose(back, c0[:, 0] * 10 + c0[:, 1], rtol=1e-12)


@pytest.mark.parametrize("field,index,value,message", [
    ("fpi", 5, 10_000, "face vertex index out of range"),
    ("fpi", 0, -1, "face vertex index out of range"),
    ("owner", 3, -1, "face owner out of range"),
    ("patch", -1, 99, "boundary face without a valid patch"),
    ("neighbour", 0, -5, "boundary face without a valid patch"),
    ("fpo", 2, 0, "face point offsets must be non-decreasing"),
    ("vertices", 4, float("nan"), "non-finite vertex"),
])
def test_arrays_are_checked_before_use(field, index, value, message):
    """bsim_mesh_from_arrays rejects what Mesh::polyhedral rejects
    (src/core/mesh.cpp:321-367) with rc 2, before any array is indexed."""
    a = {k: np.array(v, copy=True) for k, v in MeshData.hex((3, 2, 2)).arrays().items()}
    a[field][index] = value
    with pytest.raises(ValueError, match=message):
        MeshData.from_arrays(a)


def test_valid_arrays_round_trip():
    a = MeshData.hex((3, 2, 2)).arrays()
    b = MeshData.from_arrays(a).arrays()
    for k in a:
        np.testing.assert_array_equal(a[k], b[k])
