"""LinkedCellList build and permute over the C ABI
(core/src/Cabana_LinkedCellList.hpp, core/src/Cabana_Sort.hpp)."""
import ctypes

import torch

from . import _native as N
from .grid import particle_descriptor


def _arr3(v):
    return (ctypes.c_double * 3)(*[float(x) for x in v])


class LinkedCellList:
    """Cell binning of particle positions: counts, offsets and the
    permutation (absolute particle indices), identical to a Serial build."""

    def __init__(self, positions, grid_delta, grid_min, grid_max, begin=0, end=None):
        self._delta = [float(v) for v in grid_delta]
        self._min = [float(v) for v in grid_min]
        self._max = [float(v) for v in grid_max]
        nb = (ctypes.c_int32 * 3)()
        N.check(N.load().cabana_amd_linked_cell_num_bins(
            _arr3(self._delta), _arr3(self._min), _arr3(self._max), nb))
        self._nbin = [nb[0], nb[1], nb[2]]
        self._positions = positions
        dev = positions.data.device if hasattr(positions, "vector_length") else positions.device
        ncell = self._nbin[0] * self._nbin[1] * self._nbin[2]
        self.counts = torch.zeros(ncell, dtype=torch.int32, device=dev)
        self.offsets = torch.zeros(ncell, dtype=torch.int64, device=dev)
        n = particle_descriptor(positions).num_particle
        self.build(positions, begin, n if end is None else end)

    def build(self, positions, begin, end):
        self._begin = int(begin)
        self._end = int(end)
        dev = self.counts.device
        self.permutes = torch.zeros(max(self._end - self._begin, 0), dtype=torch.int64,
                                    device=dev)
        pos = particle_descriptor(positions)
        ctx = N.context()
        # A new serial (issued by the library, unique in the process) before
        # the build: counts/offsets change even when the build fails part way.
        # The library reuses the brick part tables it derived from one build
        # while the serial holds.
        self._serial = N.load().cabana_amd_new_binning_serial()
        N.check(N.load().cabana_amd_linked_cell_build(
            ctx, ctypes.byref(pos), self._begin, self._end, _arr3(self._delta),
            _arr3(self._min), _arr3(self._max), self.counts.data_ptr(),
            self.offsets.data_ptr(),
            self.permutes.data_ptr() if self.permutes.numel() else None,
            N.current_stream()), ctx)

    def numBin(self, d):
        return self._nbin[d]

    def totalBins(self):
        return self._nbin[0] * self._nbin[1] * self._nbin[2]

    def cardinalIndex(self, i, j, k):
        return (i * self._nbin[1] + j) * self._nbin[2] + k

    def binSize(self, i, j, k):
        return int(self.counts[self.cardinalIndex(i, j, k)])

    def binOffset(self, i, j, k):
        return int(self.offsets[self.cardinalIndex(i, j, k)])

    def permutation(self, i):
        return int(self.permutes[i])

    def rangeBegin(self):
        return self._begin

    def rangeEnd(self):
        return self._end

    def binning_descriptor(self):
        """Binning handle for the interpolation calls (the particles must have
        been permuted with this list)."""
        b = N.Binning()
        b.counts = self.counts.data_ptr()
        b.offsets = self.offsets.data_ptr()
        b.begin = self._begin
        b.end = self._end
        for d in range(3):
            b.num_bin[d] = self._nbin[d]
        b.serial = self._serial
        return b


def createLinkedCellList(positions, grid_delta, grid_min, grid_max, begin=0, end=None):
    return LinkedCellList(positions, grid_delta, grid_min, grid_max, begin, end)


def permute(linked_cell_list, *fields):
    """permute(linked_cell_list, slice) for one field; for several,
    permute(linked_cell_list, aosoa) over those members: each is permuted
    once, in one fused pass where their layout allows.  The members must not
    share memory (passing one field twice is rejected)."""
    ctx = N.context()
    perm = (linked_cell_list.permutes.data_ptr() if linked_cell_list.permutes.numel()
            else None)
    descs = [particle_descriptor(f) for f in fields]
    if len(descs) == 1:
        N.check(N.load().cabana_amd_permute(
            ctx, perm, linked_cell_list.rangeBegin(), linked_cell_list.rangeEnd(),
            ctypes.byref(descs[0]), N.current_stream()), ctx)
        return
    if len({d.data for d in descs}) != len(descs):
        raise ValueError("permute: the same field was passed twice")
    arr = (type(descs[0]) * len(descs))(*descs) if descs else None
    N.check(N.load().cabana_amd_permute_members(
        ctx, perm, linked_cell_list.rangeBegin(), linked_cell_list.rangeEnd(),
        arr, len(descs), N.current_stream()), ctx)
