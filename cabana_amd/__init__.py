"""cabana_amd: MI355X-native particle<->grid interpolation with Cabana's
p2g / g2p operator semantics.  The compute runs in hand-written HIP kernels
(libcabana_amd.so); this package is the host-side mirror of the reference
interface over the C ABI."""
from .grid import (  # noqa: F401
    AoSoASlice, Array, ArrayLayout, ArrayOp, BovWriter, Cell, Dim, Edge, Experimental, Face, Ghost,
    adjoint_gather,
    GlobalGrid,
    GlobalMesh, Halo,
    IndexSpace, Local, LocalGrid, LocalMesh, Node, NodeHaloPattern, Own, ScatterReduce,
    Spline, check_domain, createArray, createArrayLayout, createGlobalGrid, createHalo,
    createLocalGrid, createLocalMesh, createScalarGradientG2P, createScalarGradientP2G,
    createScalarValueG2P, createScalarValueP2G, createTensorDivergenceP2G,
    createUniformGlobalMesh, createVectorDivergenceG2P, createVectorDivergenceP2G,
    createVectorGradientG2P, createVectorValueG2P, createVectorValueP2G, g2p, g2p2g, g2p_adjoint,
    grid_descriptor, p2g, p2g_adjoint, particleMigrate, particle_descriptor, profile_enable,
    profile_report, profile_reset, set_auto_binning, set_deterministic)
from .linkedcell import LinkedCellList, createLinkedCellList, permute  # noqa: F401
from . import autograd  # noqa: F401,E402
