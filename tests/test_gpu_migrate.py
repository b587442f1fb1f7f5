"""particleMigrate on one rank (cabana_amd_particle_migrate) against the
oracle, bit for bit: the reference's test geometry
(grid/unit_test/tstParticleDistributor3d.hpp), random clouds reaching 0.6
extents past every face with mixed periodicity, members of every layout
(row-major, column-major, AoSoA slices, int32, float32 tensors) compacted in
order, the no-op path below the band and more particles than one scan
block holds."""
import numpy as np
import pytest
import torch

import cabana_amd as ca
from test_oracle_migrate import CELL, LOW, NCELL, ghost_cell_centres

pytestmark = pytest.mark.gpu


def local_grid(halo, periodic):
    high = LOW + CELL * NCELL
    gm = ca.createUniformGlobalMesh(LOW, high, CELL)
    return ca.createLocalGrid(ca.createGlobalGrid(gm, list(periodic)), halo)


def omesh(oracle, halo, periodic):
    return oracle.OMesh(LOW, LOW + CELL * NCELL, CELL, halo=halo, periodic=tuple(periodic))


@pytest.mark.parametrize("periodic", [False, True])
@pytest.mark.parametrize("data_halo,test_halo,force", [(0, 1, False), (1, 1, False),
                                                       (2, 2, False), (2, 0, False),
                                                       (2, 2, True), (1, 0, True)])
def test_reference_geometry_matches_oracle(oracle, periodic, data_halo, test_halo, force):
    per = (periodic,) * 3
    x = ghost_cell_centres(data_halo, periodic)
    ids = np.arange(len(x), dtype=np.int32)
    xt = torch.tensor(x, device="cuda")
    it = torch.tensor(ids, device="cuda")
    migrated, n = ca.particleMigrate(local_grid(data_halo, per), xt, [xt, it], test_halo,
                                     force_migrate=force)
    om, y, keep = oracle.particle_migrate(omesh(oracle, data_halo, per), x, test_halo, force)
    assert migrated == om
    if keep is None:
        keep = np.ones(len(x), dtype=bool)
    assert n == int(keep.sum())
    np.testing.assert_array_equal(xt[:n].cpu().numpy(), y[keep])
    np.testing.assert_array_equal(it[:n].cpu().numpy(), ids[keep])


def _aosoa(values, vl):
    """AoSoA of two members, double[3] and double, vector length vl: returns
    the storage and the two slices."""
    n = values.shape[0]
    nblk = (n + vl - 1) // vl
    stride = vl * 4
    data = torch.zeros(nblk * stride, dtype=torch.float64, device="cuda")
    s0 = ca.AoSoASlice(data, n, (3,), vl, stride)
    s1 = ca.AoSoASlice(data[3 * vl:], n, (), vl, stride)
    s0.from_tensor(torch.tensor(values[:, :3], device="cuda"))
    s1.from_tensor(torch.tensor(values[:, 3], device="cuda"))
    return data, s0, s1


@pytest.mark.parametrize("periodic", [(True, False, True), (True, True, True),
                                      (False, False, False)])
@pytest.mark.parametrize("n", [1, 1000, 300_001])
def test_random_cloud_matches_oracle(oracle, periodic, n):
    rng = np.random.default_rng(n + sum(periodic))
    high = LOW + CELL * NCELL
    x = LOW + rng.uniform(-0.6, 1.6, size=(n, 3)) * (high - LOW)
    if n > 2:
        x[0] = high
        x[1] = LOW
    ids = np.arange(n, dtype=np.int32)
    f32 = rng.normal(size=(n, 3, 3)).astype(np.float32)
    col = rng.normal(size=(n, 2))
    extra = rng.normal(size=(n, 4))
    xt = torch.tensor(x, device="cuda")
    it = torch.tensor(ids, device="cuda")
    ft = torch.tensor(f32, device="cuda")
    ct = torch.tensor(col.T.copy(), device="cuda").T          # column-major (n, 2)
    _, s0, s1 = _aosoa(extra, 16)
    migrated, count = ca.particleMigrate(local_grid(1, periodic), xt, [xt, it, ft, ct, s0, s1],
                                         1)
    om, y, keep = oracle.particle_migrate(omesh(oracle, 1, periodic), x, 1)
    assert migrated == om
    if keep is None:
        keep = np.ones(n, dtype=bool)
    assert count == int(keep.sum())
    np.testing.assert_array_equal(xt[:count].cpu().numpy(), y[keep])
    np.testing.assert_array_equal(it[:count].cpu().numpy(), ids[keep])
    np.testing.assert_array_equal(ft[:count].cpu().numpy(), f32[keep])
    np.testing.assert_array_equal(ct[:count].cpu().numpy(), col[keep])
    got = torch.cat([s0.to_tensor(), s1.to_tensor()[:, None]], 1)[:count].cpu().numpy()
    np.testing.assert_array_equal(got, extra[keep])


def test_positions_outside_the_members_are_shifted_not_compacted(oracle):
    """positions is a separate view, not a member of the particle set: it is
    shifted in place and keeps its length; the members are compacted."""
    rng = np.random.default_rng(8)
    high = LOW + CELL * NCELL
    per = (False, True, True)
    x = LOW + rng.uniform(-0.3, 1.3, size=(5000, 3)) * (high - LOW)
    ids = np.arange(len(x), dtype=np.int32)
    xt = torch.tensor(x, device="cuda")
    it = torch.tensor(ids, device="cuda")
    migrated, n = ca.particleMigrate(local_grid(2, per), xt, [it], 2)
    om, y, keep = oracle.particle_migrate(omesh(oracle, 2, per), x, 2)
    assert migrated and om
    np.testing.assert_array_equal(xt.cpu().numpy(), y)
    np.testing.assert_array_equal(it[:n].cpu().numpy(), ids[keep])


def test_no_particle_outside_the_band_is_a_no_op(oracle):
    rng = np.random.default_rng(9)
    x = LOW + rng.uniform(0.01, 0.99, size=(20_000, 3)) * (CELL * NCELL)
    xt = torch.tensor(x, device="cuda")
    migrated, n = ca.particleMigrate(local_grid(1, (True,) * 3), xt, [xt], 1)
    assert not migrated and n == len(x)
    np.testing.assert_array_equal(xt.cpu().numpy(), x)
    migrated, n = ca.particleMigrate(local_grid(1, (True,) * 3), xt, [xt], 1,
                                     force_migrate=True)
    assert migrated and n == len(x)
    np.testing.assert_array_equal(xt.cpu().numpy(), x)


def test_empty_and_fully_removed_sets():
    per = (False,) * 3
    xt = torch.zeros(0, 3, dtype=torch.float64, device="cuda")
    assert ca.particleMigrate(local_grid(1, per), xt, [xt], 0, force_migrate=True) == (True, 0)
    assert ca.particleMigrate(local_grid(1, per), xt, [xt], 0) == (False, 0)
    x = np.tile(LOW - 0.5 * CELL, (3000, 1))
    xt = torch.tensor(x, device="cuda")
    assert ca.particleMigrate(local_grid(1, per), xt, [xt], 0, force_migrate=True) == (True, 0)
