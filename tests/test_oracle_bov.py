"""Oracle BovWriter::writeTimeStep on one rank against the reference's own
test (grid/unit_test/tstBovWriter.hpp writeTest3d :37-200: a 22 x 19 x 21
periodic grid, cell size 0.23 from the cell counts, halo 0, a scalar cell
field and a 3-dof node field read back in KJI order) and the header the
writer prints (grid/src/Cabana_Grid_BovWriter.hpp:276-337)."""
import numpy as np

LOW = np.array([1.2, 3.3, -2.8])
NCELL = np.array([22, 19, 21])


def reference_case(oracle, halo=0):
    high = LOW + 0.23 * NCELL
    cell = [(high[d] - LOW[d]) / NCELL[d] for d in range(3)]
    return oracle.OMesh(LOW, high, cell, halo=halo), high


def fields(mesh):
    """writeTest3d :66-110: cell (i, j, k) holds 1 + |cos(2 pi i/N0) cos(2 pi
    j/N1) cos(2 pi k/N2)|; owned node (i, j, k) holds (1 + |cos(2 pi i/N0)|,
    ...)."""
    pi2 = 8.0 * np.arctan(1.0)
    h = mesh.halo
    n = NCELL
    args = [np.arange(n[d]) / n[d] for d in range(3)]
    cell = np.zeros(tuple(n + 2 * h) + (1,))
    c = [np.cos(pi2 * a) for a in args]
    cell[h:h + n[0], h:h + n[1], h:h + n[2], 0] = 1.0 + np.abs(
        c[0][:, None, None] * c[1][None, :, None] * c[2][None, None, :])
    node = np.zeros(tuple(n + 2 * h + 1) + (3,))
    for d in range(3):
        shape = [1, 1, 1]
        shape[d] = n[d]
        node[h:h + n[0], h:h + n[1], h:h + n[2], d] = (1.0 + np.abs(c[d])).reshape(shape)
    return cell, node


def test_reference_write_test(oracle, tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    mesh, _ = reference_case(oracle)
    cell, node = fields(mesh)
    oracle.halo_gather(mesh, node, 0)
    oracle.bov_write(mesh, cell, "cell", "cell_field_3d", "grid_cell_field_3d", 302, 3.43)
    oracle.bov_write(mesh, node, "node", "node_field_3d", "grid_node_field_3d", 1972, 12.457)
    pi2 = 8.0 * np.arctan(1.0)
    # :125-153: the cell file in KJI order.
    data = np.fromfile("grid_cell_field_3d_000302.dat", dtype=np.float64)
    cid = 0
    for k in range(NCELL[2]):
        for j in range(NCELL[1]):
            for i in range(NCELL[0]):
                want = 1.0 + abs(np.cos(pi2 * i / NCELL[0]) * np.cos(pi2 * j / NCELL[1]) *
                                 np.cos(pi2 * k / NCELL[2]))
                assert np.isclose(data[cid], want, rtol=4e-7, atol=0)
                cid += 1
    assert cid == data.size
    # :155-197: the node file, three values per node, NCELL + 1 nodes per
    # dimension (the closing node repeats node 0 after the gather).
    data = np.fromfile("grid_node_field_3d_001972.dat", dtype=np.float64).reshape(
        NCELL[2] + 1, NCELL[1] + 1, NCELL[0] + 1, 3)
    for d, axis in ((0, 2), (1, 1), (2, 0)):
        idx = np.arange(NCELL[d] + 1)
        want = 1.0 + np.abs(np.cos(pi2 * idx / NCELL[d]))
        shape = [1, 1, 1]
        shape[axis] = NCELL[d] + 1
        np.testing.assert_allclose(data[..., d], np.broadcast_to(want.reshape(shape),
                                                                 data.shape[:3]),
                                   rtol=4e-7, atol=0)


def test_header_text(oracle, tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    mesh, _ = reference_case(oracle)
    cell, node = fields(mesh)
    oracle.bov_write(mesh, node, "node", "node_field_3d", "grid_node_field_3d", 1972, 12.457)
    oracle.bov_write(mesh, cell, "cell", "cell_field_3d", "grid_cell_field_3d", 302, 3.43)
    assert open("grid_node_field_3d_001972.bov").read() == (
        "TIME: 12.457\n"
        "DATA_FILE: grid_node_field_3d_001972.dat\n"
        "DATA_SIZE: 23 20 22 \n"
        "DATA_FORMAT: DOUBLE\n"
        "VARIABLE: node_field_3d\n"
        "DATA_ENDIAN: LITTLE\n"
        "CENTERING: nodal\n"
        "BRICK_ORIGIN: 1.2 3.3 -2.8 \n"
        "BRICK_SIZE: 5.06 4.37 4.83 \n"
        "DATA_COMPONENTS: 3\n")
    txt = open("grid_cell_field_3d_000302.bov").read().splitlines()
    assert txt[0] == "TIME: 3.43" and txt[2] == "DATA_SIZE: 22 19 21 "
    assert txt[6] == "CENTERING: zonal" and txt[9] == "DATA_COMPONENTS: 1"


def test_writer_gathers_periodic_closing_nodes(oracle, tmp_path, monkeypatch):
    """Without the caller's gather, writeTimeStep's own width-0 gather
    (:227-240) fills the closing nodes; a non-periodic dimension keeps its
    owned closing node."""
    monkeypatch.chdir(tmp_path)
    high = LOW + 0.23 * NCELL
    rng = np.random.default_rng(2)
    for periodic in ((True, True, True), (True, False, True)):
        mesh = oracle.OMesh(LOW, high, 0.23, halo=2, periodic=periodic)
        g = rng.normal(size=mesh.num_node + (2,))
        ref = g.copy()
        oracle.halo_gather(mesh, ref, 0)
        oracle.bov_write(mesh, g, "node", "f", "out", 7, 0.5)
        data = np.fromfile("out_000007.dat").reshape(NCELL[2] + 1, NCELL[1] + 1,
                                                      NCELL[0] + 1, 2)
        own = ref[2:3 + NCELL[0], 2:3 + NCELL[1], 2:3 + NCELL[2]]
        np.testing.assert_array_equal(data, own.transpose(2, 1, 0, 3))
        np.testing.assert_array_equal(g, ref)
        if periodic[1]:
            np.testing.assert_array_equal(data[:, -1], data[:, 0])
