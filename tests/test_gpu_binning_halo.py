"""HIP binning (LinkedCellList build + permute) bit-exact against the
oracle's Serial build, and the HIP node halo against the oracle."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def test_cartesian_and_linked_cell_known_answers(ca, oracle):
    # tstLinkedCellList.hpp:30-104, 281-365 on the GPU: 10^3 cell centres.
    nx = 10
    n = nx ** 3
    pos = np.zeros((n, 3))
    ids = np.zeros((n, 3), dtype=np.int32)
    for k in range(nx):
        for j in range(nx):
            for i in range(nx):
                pid = i + j * nx + k * nx * nx
                ids[pid] = [i, j, k]
                pos[pid] = [i + 0.5, j + 0.5, k + 0.5]
    for begin, end in ((0, n), (250, 750)):
        tpos = torch.tensor(pos, device="cuda")
        tids = torch.tensor(ids, device="cuda")
        lcl = ca.createLinkedCellList(tpos, [1.0] * 3, [0.0] * 3, [10.0] * 3, begin, end)
        ca.permute(lcl, tids)
        got = tids.cpu().numpy()
        counts = lcl.counts.cpu().numpy()
        offsets = lcl.offsets.cpu().numpy()
        particle_id = 0
        for i in range(nx):
            for j in range(nx):
                for k in range(nx):
                    original = i + j * nx + k * nx * nx
                    cell = (i * nx + j) * nx + k
                    if begin <= original < end:
                        assert tuple(got[begin + particle_id]) == (i, j, k)
                        assert counts[cell] == 1 and offsets[cell] == particle_id
                        particle_id += 1
                    else:
                        assert counts[cell] == 0
        for pid in list(range(0, begin)) + list(range(end, n)):
            assert tuple(got[pid]) == (pid % nx, (pid // nx) % nx, pid // (nx * nx))


@pytest.mark.parametrize("n,ppc", [(20000, 8), (150000, 40)])
def test_linked_cell_bit_exact(ca, oracle, n, ppc):
    """Several particles per bin (uniform and clustered): counts, offsets and
    the permutation equal the Serial build exactly."""
    rng = np.random.default_rng(n)
    lo = np.array([-1.0, 0.5, 2.0])
    ncell = max(1, int(round((n / ppc) ** (1 / 3))))
    hi = lo + 0.37 * ncell
    if ppc > 20:
        centres = lo + rng.uniform(0.2, 0.8, size=(4, 3)) * (hi - lo)
        pos = centres[rng.integers(0, 4, n)] + rng.normal(scale=0.3, size=(n, 3))
        pos = np.clip(pos, lo, hi)
    else:
        pos = lo + rng.uniform(0, 1, size=(n, 3)) * (hi - lo)
    # Points exactly on cell faces and on the upper boundary.
    pos[:50, 0] = lo[0] + 0.37 * rng.integers(0, ncell + 1, 50)
    delta = [0.37] * 3
    counts, offsets, perm, _ = oracle.lcl_build(pos, delta, lo, hi)
    lcl = ca.createLinkedCellList(torch.tensor(pos, device="cuda"), delta, lo, hi)
    assert np.array_equal(lcl.counts.cpu().numpy(), counts)
    assert np.array_equal(lcl.offsets.cpu().numpy(), offsets)
    assert np.array_equal(lcl.permutes.cpu().numpy(), perm)


@pytest.mark.parametrize("case", ["three_passes", "crowded_bin", "subrange"])
def test_linked_cell_sort_shapes(ca, oracle, case):
    """Bin counts that need three radix passes (2^18 bins), one bin holding
    most particles, and a begin/end subrange: bit-exact Serial build."""
    rng = np.random.default_rng({"three_passes": 1, "crowded_bin": 2, "subrange": 3}[case])
    nx = 64
    lo, hi = np.zeros(3), np.full(3, float(nx))
    n = 300001
    pos = rng.uniform(0, nx, size=(n, 3))
    if case == "crowded_bin":
        pos[: 2 * n // 3] = 17.0 + rng.uniform(0, 1, size=(2 * n // 3, 3))
    begin, end = (0, n) if case != "subrange" else (12345, n - 777)
    counts, offsets, perm, _ = oracle.lcl_build(pos, [1.0] * 3, lo, hi, begin=begin, end=end)
    lcl = ca.createLinkedCellList(torch.tensor(pos, device="cuda"), [1.0] * 3, lo, hi,
                                  begin, end)
    assert np.array_equal(lcl.counts.cpu().numpy(), counts)
    assert np.array_equal(lcl.offsets.cpu().numpy(), offsets)
    assert np.array_equal(lcl.permutes.cpu().numpy(), perm)


@pytest.mark.parametrize("case", ["empty_range", "single_bin", "four_passes"])
def test_linked_cell_sort_edges(ca, oracle, case):
    """An empty [begin, end), a grid of one bin (one-bit keys) and a grid of
    more than 2^24 bins (four radix passes): counts, offsets and permutes
    equal the Serial build."""
    rng = np.random.default_rng({"empty_range": 4, "single_bin": 5, "four_passes": 6}[case])
    if case == "four_passes":
        nx = (256, 256, 257)
        n = 100000
    else:
        nx = (1, 1, 1) if case == "single_bin" else (4, 4, 4)
        n = 5000
    lo = np.zeros(3)
    hi = np.array(nx, dtype=np.float64)
    pos = rng.uniform(0, 1, size=(n, 3)) * hi
    begin, end = (2500, 2500) if case == "empty_range" else (0, n)
    counts, offsets, perm, _ = oracle.lcl_build(pos, [1.0] * 3, lo, hi, begin=begin, end=end)
    lcl = ca.createLinkedCellList(torch.tensor(pos, device="cuda"), [1.0] * 3, lo, hi,
                                  begin, end)
    assert np.array_equal(lcl.counts.cpu().numpy(), counts)
    assert np.array_equal(lcl.offsets.cpu().numpy(), offsets)
    assert np.array_equal(lcl.permutes.cpu().numpy()[:end - begin], perm[:end - begin])


@pytest.mark.parametrize("begin_end", [(0, None), (777, -1234)])
def test_permute_row_layouts(ca, oracle, begin_end):
    """permute of row-contiguous fields (FP64 (n,3), FP32 (n,3,3), int32 (n,))
    and of a padded-row view ((n,4)[:, :3]) equals indexing by the Serial
    permutation; entries outside [begin, end) are untouched."""
    rng = np.random.default_rng(31)
    n = 50001
    begin, end = begin_end[0], (n if begin_end[1] is None else n + begin_end[1])
    pos = rng.uniform(0, 6.0, size=(n, 3))
    counts, offsets, perm, _ = oracle.lcl_build(pos, [0.5] * 3, [0.0] * 3, [6.0] * 3,
                                                begin=begin, end=end)
    lcl = ca.createLinkedCellList(torch.tensor(pos, device="cuda"), [0.5] * 3, [0.0] * 3,
                                  [6.0] * 3, begin, end)
    assert np.array_equal(lcl.permutes.cpu().numpy(), perm)
    a64 = rng.normal(size=(n, 3))
    a32 = rng.normal(size=(n, 3, 3)).astype(np.float32)
    ai = rng.integers(-2**31, 2**31 - 1, size=n, dtype=np.int32)
    wide = rng.normal(size=(n, 4)).astype(np.float32)
    t64 = torch.tensor(a64, device="cuda")
    t32 = torch.tensor(a32, device="cuda")
    ti = torch.tensor(ai, device="cuda")
    twide = torch.tensor(wide, device="cuda")
    ca.permute(lcl, t64, t32, ti, twide[:, :3])

    def expect(x):
        y = x.copy()
        y[begin:end] = x[perm]
        return y

    wide_ref = wide.copy()
    wide_ref[:, :3] = expect(wide[:, :3])
    assert np.array_equal(t64.cpu().numpy(), expect(a64))
    assert np.array_equal(t32.cpu().numpy(), expect(a32))
    assert np.array_equal(ti.cpu().numpy(), expect(ai))
    assert np.array_equal(twide.cpu().numpy(), wide_ref)


@pytest.mark.parametrize("begin_end", [(0, None), (513, -700)])
def test_permute_members_fused_and_single(ca, oracle, begin_end):
    """permute(binning, aosoa) over 19 members: 18 row-contiguous ones (more
    than one fused pass holds, mixed FP64/FP32/int32/int64 and word counts,
    one padded-row view) and one AoSoA member (VL = 32), all bitwise equal to
    indexing by the Serial permutation; the same field twice is rejected."""
    rng = np.random.default_rng(47)
    n = 20011
    begin, end = begin_end[0], (n if begin_end[1] is None else n + begin_end[1])
    pos = rng.uniform(0, 5.0, size=(n, 3))
    _, _, perm, _ = oracle.lcl_build(pos, [0.5] * 3, [0.0] * 3, [5.0] * 3,
                                     begin=begin, end=end)
    lcl = ca.createLinkedCellList(torch.tensor(pos, device="cuda"), [0.5] * 3, [0.0] * 3,
                                  [5.0] * 3, begin, end)
    shapes = [((), np.float64), ((3,), np.float32), ((3, 3), np.float32), ((), np.int32),
              ((2,), np.int64), ((4,), np.float64), ((5,), np.int32), ((3,), np.float64)]
    arrays = []
    for k in range(17):
        shp, dt = shapes[k % len(shapes)]
        if np.issubdtype(dt, np.integer):
            a = rng.integers(-2**31, 2**31 - 1, size=(n,) + shp).astype(dt)
        else:
            a = rng.normal(size=(n,) + shp).astype(dt)
        arrays.append(a)
    wide = rng.normal(size=(n, 6)).astype(np.float32)
    tensors = [torch.tensor(a, device="cuda") for a in arrays]
    twide = torch.tensor(wide, device="cuda")
    vl = 32
    nsoa = (n + vl - 1) // vl
    raw = torch.tensor(rng.normal(size=nsoa * vl * 4), device="cuda")
    member = ca.AoSoASlice(raw[vl:], n, (3,), vl, vl * 4)
    m0 = member.to_tensor().cpu().numpy()
    ca.permute(lcl, *tensors, twide[:, 1:4], member)

    def expect(x):
        y = x.copy()
        y[begin:end] = x[perm]
        return y

    for a, t in zip(arrays, tensors):
        assert np.array_equal(t.cpu().numpy(), expect(a))
    wide_ref = wide.copy()
    wide_ref[:, 1:4] = expect(wide[:, 1:4])
    assert np.array_equal(twide.cpu().numpy(), wide_ref)
    assert np.array_equal(member.to_tensor().cpu().numpy(), expect(m0))
    with pytest.raises(ValueError):
        ca.permute(lcl, tensors[0], tensors[0])
    ca.permute(lcl)


def test_linked_cell_aosoa_positions_and_permute(ca, oracle):
    """Positions and a tensor member in one AoSoA (VL = 64): build from the
    slice, permute the slice, compare bitwise with the oracle."""
    rng = np.random.default_rng(8)
    n, vl = 3001, 64
    nsoa = (n + vl - 1) // vl
    stride = vl * (3 + 9)
    raw = torch.tensor(rng.normal(size=nsoa * stride), device="cuda")
    pos = ca.AoSoASlice(raw, n, (3,), vl, stride)
    ten = ca.AoSoASlice(raw[3 * vl:], n, (3, 3), vl, stride)
    p = rng.uniform(0, 4.0, size=(n, 3))
    pos.from_tensor(torch.tensor(p, device="cuda"))
    raw_np = raw.cpu().numpy().copy()
    lcl = ca.createLinkedCellList(pos, [0.5] * 3, [0.0] * 3, [4.0] * 3)
    ca.permute(lcl, pos, ten)
    counts, offsets, perm, _ = oracle.lcl_build(raw_np, [0.5] * 3, [0.0] * 3, [4.0] * 3,
                                                vl=vl, soa_stride=stride, n=n)
    assert np.array_equal(lcl.permutes.cpu().numpy(), perm)
    f_pos = oracle.pfield(raw_np, vl=vl, soa_stride=stride, ncomp=3, n=n)
    oracle.permute(perm, 0, n, f_pos)
    tview = raw_np[3 * vl:].copy()
    f_ten = oracle.pfield(tview, vl=vl, soa_stride=stride, ncomp=9, n=n)
    oracle.permute(perm, 0, n, f_ten)
    raw_np[3 * vl:] = tview
    assert np.array_equal(raw.cpu().numpy(), raw_np)


def test_linked_cell_out_of_grid(ca, oracle):
    pos = torch.tensor([[0.5, 0.5, 0.5], [5.0, 0.5, 0.5]], device="cuda")
    from cabana_amd._native import CabanaAmdError
    with pytest.raises(CabanaAmdError):
        ca.createLinkedCellList(pos, [1.0] * 3, [0.0] * 3, [2.0] * 3)


@pytest.mark.parametrize("width", [1, 2, 3])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_halo_pattern(ca, oracle, width, dtype):
    # tstHalo3d.hpp:181-230 geometry with a node array (4 dofs, halo 3).
    cell = 0.23
    low = np.array([1.2, 3.3, -2.8])
    high = low + cell * np.array([17, 20, 21])
    gm = ca.createUniformGlobalMesh(low, high, cell)
    lg = ca.createLocalGrid(ca.createGlobalGrid(gm, [True] * 3), 3)
    arr = ca.createArray("a", ca.createArrayLayout(lg, 4, ca.Node()), dtype=dtype)
    ca.ArrayOp.assign(arr, 0.0, ca.Ghost())
    ca.ArrayOp.assign(arr, 1.0, ca.Own())
    halo = ca.createHalo(ca.NodeHaloPattern(), width, arr)
    halo.gather(arr)
    mesh = oracle.OMesh(low, high, cell, halo=3)
    ref = mesh.grid(4)
    ref[mesh.owned_slices()] = 1.0
    oracle.halo_gather(mesh, ref, width)
    assert np.array_equal(arr.view().cpu().numpy().astype(np.float64), ref)
    halo.scatter(ca.ScatterReduce.Sum, arr)
    oracle.halo_scatter_sum(mesh, ref, width)
    assert np.array_equal(arr.view().cpu().numpy().astype(np.float64), ref)


@pytest.mark.parametrize("periodic", [(True, False, True), (False, False, False)])
def test_halo_random_values_mixed_periodicity(ca, oracle, periodic):
    rng = np.random.default_rng(1)
    cell = 0.5
    low = np.array([0.0, 1.0, -2.0])
    high = low + cell * np.array([6, 7, 5])
    gm = ca.createUniformGlobalMesh(low, high, cell)
    lg = ca.createLocalGrid(ca.createGlobalGrid(gm, periodic), 2)
    arr = ca.createArray("a", ca.createArrayLayout(lg, 3, ca.Node()), memory_layout="left")
    vals = rng.normal(size=arr.view().shape)
    arr.view().copy_(torch.tensor(vals))
    mesh = oracle.OMesh(low, high, cell, halo=2, periodic=periodic)
    ref = np.ascontiguousarray(vals)
    halo = ca.createHalo(ca.NodeHaloPattern(), -1, arr)
    halo.scatter(ca.ScatterReduce.Sum, arr)
    oracle.halo_scatter_sum(mesh, ref)
    np.testing.assert_allclose(arr.view().cpu().numpy(), ref, rtol=1e-14, atol=1e-14)
    halo.gather(arr)
    oracle.halo_gather(mesh, ref)
    np.testing.assert_allclose(arr.view().cpu().numpy(), ref, rtol=1e-14, atol=1e-14)
