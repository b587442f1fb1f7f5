"""Oracle particleMigrate on one rank against the reference's own test cases
(grid/unit_test/tstParticleDistributor3d.hpp: migrateTest :41-176,
localOnlyTest :178-257, removeOutsideTest :259-326, on the 18 x 15 x 9 cell,
0.23 cell-size grid of createGrid :328-343) with a single block, where every
neighbour of a periodic dimension is the rank itself."""
import itertools

import numpy as np
import pytest

LOW = np.array([1.2, 3.3, -2.8])
NCELL = np.array([18, 15, 9])
CELL = 0.23


def mesh(oracle, halo, periodic):
    high = LOW + CELL * NCELL
    return oracle.OMesh(LOW, high, CELL, halo=halo, periodic=(periodic,) * 3)


def ghost_cell_centres(halo, periodic):
    """migrateTest :66-97: a particle at the centre of every ghosted cell
    whose neighbour rank exists (every cell when periodic, the owned cells
    otherwise on one rank)."""
    ghost_low = LOW - halo * CELL
    lo = 0 if periodic else halo
    ranges = [range(lo, (NCELL[d] + 2 * halo) - lo) for d in range(3)]
    ijk = np.array(list(itertools.product(*ranges)), dtype=np.float64)
    return ghost_low + (ijk + 0.5) * CELL


@pytest.mark.parametrize("periodic", [False, True])
@pytest.mark.parametrize("data_halo", [0, 1, 2])
@pytest.mark.parametrize("test_halo", [0, 1, 2])
def test_migrate_known_answers(oracle, periodic, data_halo, test_halo):
    """migrateTest :131-175: with test_halo > 0 (something lies outside the
    band) every particle ends inside the owned domain and none is lost; with
    test_halo = 0 nothing moves."""
    m = mesh(oracle, data_halo, periodic)
    x = ghost_cell_centres(data_halo, periodic)
    migrated, y, keep = oracle.particle_migrate(m, x, test_halo)
    own_lo = LOW + CELL * 0
    own_hi = LOW + CELL * NCELL
    if test_halo > 0:
        # Periodic: the ghost-cell particles lie outside the band, so the
        # migration runs.  Non-periodic: only owned cells hold particles and
        # the band may contain them all; either way all stay, inside.
        assert migrated or not periodic
        assert keep is None or keep.all()
        assert (y >= own_lo).all() and (y <= own_hi).all()
    else:
        assert not migrated
        np.testing.assert_array_equal(y, x)


def test_forced_migration_keeps_owned_particles(oracle):
    """localOnlyTest :178-257 (and migrateTest with force_comm, :370-374):
    particles at owned cell centres with a halo of 2 all stay, unmoved."""
    for periodic in (False, True):
        m = mesh(oracle, 2, periodic)
        x = ghost_cell_centres(2, False)
        migrated, y, keep = oracle.particle_migrate(m, x, 0, force=True)
        assert migrated and keep.all()
        np.testing.assert_array_equal(y, x)


def test_particles_outside_a_non_periodic_domain_are_removed(oracle):
    """removeOutsideTest :259-326: cell centres of the ghost layer beyond each
    face of a non-periodic domain (halo 2) are all removed."""
    m = mesh(oracle, 2, False)
    ghost_low = LOW - 2 * CELL
    pts = []
    for d in range(3):
        for side in (-1, 1):
            ranges = [range(2, 2 + NCELL[e]) for e in range(3)]
            ranges[d] = range(0, 2) if side < 0 else range(2 + NCELL[d], 4 + NCELL[d])
            pts += list(itertools.product(*ranges))
    x = ghost_low + (np.array(pts, dtype=np.float64) + 0.5) * CELL
    migrated, _, keep = oracle.particle_migrate(m, x, 0, force=True)
    assert migrated and not keep.any()


def test_periodic_shift_is_one_extent(oracle):
    """Coordinates past the periodic box move by exactly one extent
    (getMigrateDestinations :135-145); inside the box (the high corner
    included) they do not move; mixed periodicity removes only particles
    leaving through the non-periodic dimension."""
    rng = np.random.default_rng(4)
    high = LOW + CELL * NCELL
    m = oracle.OMesh(LOW, high, CELL, halo=1, periodic=(True, False, True))
    x = LOW + rng.uniform(-0.6, 1.6, size=(5000, 3)) * (high - LOW)
    x[0] = high
    x[1] = LOW
    migrated, y, keep = oracle.particle_migrate(m, x, 1)
    ext = high - LOW
    assert migrated
    for d in (0, 2):
        up = x[:, d] < LOW[d]
        down = x[:, d] > high[d]
        np.testing.assert_array_equal(y[up, d], x[up, d] + ext[d])
        np.testing.assert_array_equal(y[down, d], x[down, d] - ext[d])
        np.testing.assert_array_equal(y[~(up | down), d], x[~(up | down), d])
    np.testing.assert_array_equal(y[:, 1], x[:, 1])
    out_y = (x[:, 1] < LOW[1]) | (x[:, 1] > high[1])
    np.testing.assert_array_equal(keep, ~out_y)
    assert keep[0] and keep[1]
