"""Oracle binning against the reference's known answers
(core/unit_test/tstCartesianGrid.cpp, core/unit_test/tstLinkedCellList.hpp)
and against a brute-force stable sort."""
import numpy as np
import pytest


def test_cartesian_grid_known_answers(oracle):
    # tstCartesianGrid.cpp:22-59
    cg = oracle.cart_grid([-1.0, -0.5, -0.6], [2.5, 1.5, 1.9], [0.5, 0.125, 0.25])
    assert tuple(cg.nx) == (7, 16, 10)
    assert oracle.cart_locate(cg, [-0.9, 1.4, 0.1]) == (0, 15, 2)
    # Upper boundary clamps to the last cell.
    assert oracle.cart_locate(cg, [2.5, 1.5, 1.9]) == (6, 15, 9)


def lcl_particles(nx=10, dx=1.0, x_min=0.0):
    """tstLinkedCellList.hpp:30-104: one particle per cell centre, created
    i-fastest (the reverse of the sort order)."""
    n = nx ** 3
    pos = np.zeros((n, 3))
    ids = np.zeros((n, 3))
    for k in range(nx):
        for j in range(nx):
            for i in range(nx):
                pid = i + j * nx + k * nx * nx
                ids[pid] = [i, j, k]
                pos[pid] = [x_min + (i + 0.5) * dx, x_min + (j + 0.5) * dx,
                            x_min + (k + 0.5) * dx]
    return pos, ids


@pytest.mark.parametrize("begin,end", [(0, 1000), (250, 750)])
def test_linked_cell_known_answers(oracle, begin, end):
    # tstLinkedCellList.hpp:281-365 (checkLinkedCell)
    nx = 10
    pos, ids = lcl_particles(nx)
    counts, offsets, perm, nbin = oracle.lcl_build(pos, [1.0] * 3, [0.0] * 3, [10.0] * 3,
                                                   begin, end)
    assert nbin == (nx, nx, nx)
    f = oracle.pfield(ids)
    oracle.permute(perm, begin, end, f)
    particle_id = 0
    for i in range(nx):
        for j in range(nx):
            for k in range(nx):
                original = i + j * nx + k * nx * nx
                cell = (i * nx + j) * nx + k
                if begin <= original < end:
                    sort_id = begin + particle_id
                    assert tuple(ids[sort_id]) == (i, j, k)
                    assert counts[cell] == 1
                    assert offsets[cell] == particle_id
                    particle_id += 1
                else:
                    assert counts[cell] == 0
    # Particles outside the range keep their place.
    for pid in list(range(0, begin)) + list(range(end, 1000)):
        i, j, k = pid % nx, (pid // nx) % nx, pid // (nx * nx)
        assert tuple(ids[pid]) == (i, j, k)


def test_linked_cell_is_stable_bruteforce(oracle):
    """Several particles per bin: the Serial build is the stable sort by cell
    (ascending particle index inside a bin)."""
    rng = np.random.default_rng(4)
    n = 5000
    lo = np.array([-1.0, 0.5, 2.0])
    hi = lo + np.array([3.0, 2.5, 4.0])
    pos = lo + rng.uniform(0, 1, size=(n, 3)) * (hi - lo)
    delta = [0.5, 0.5, 0.5]
    counts, offsets, perm, nbin = oracle.lcl_build(pos, delta, lo, hi)
    cg = oracle.cart_grid(lo, hi, delta)
    cells = np.array([
        (lambda ijk: (ijk[0] * nbin[1] + ijk[1]) * nbin[2] + ijk[2])(
            oracle.cart_locate(cg, p)) for p in pos])
    expect = np.argsort(cells, kind="stable")
    np.testing.assert_array_equal(perm, expect)
    np.testing.assert_array_equal(counts, np.bincount(cells, minlength=len(counts)))
    np.testing.assert_array_equal(offsets, np.concatenate([[0], np.cumsum(counts)[:-1]]))


def test_permute_aosoa_slice(oracle):
    """permute(binning, slice) through AoSoA addressing (VL = 16, a second
    member in the same SoA)."""
    rng = np.random.default_rng(2)
    n, vl, nc = 100, 16, 3
    nsoa = (n + vl - 1) // vl
    stride = vl * (nc + 2)   # this member plus a 2-wide neighbour member
    raw = rng.normal(size=nsoa * stride)
    f = oracle.pfield(raw, vl=vl, soa_stride=stride, ncomp=nc, n=n)

    def get(arr, p, c):
        return arr[(p // vl) * stride + p % vl + vl * c]

    before = np.array([[get(raw, p, c) for c in range(nc)] for p in range(n)])
    other = raw.copy()
    perm = rng.permutation(np.arange(10, 90)).astype(np.int64)
    oracle.permute(perm, 10, 90, f)
    for i in range(n):
        for c in range(nc):
            src = perm[i - 10] if 10 <= i < 90 else i
            assert get(raw, i, c) == before[src, c]
    # The neighbouring member is untouched.
    for p in range(n):
        for c in range(nc, nc + 2):
            assert get(raw, p, c) == get(other, p, c)
