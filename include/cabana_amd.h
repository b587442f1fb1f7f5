/*
 * cabana_amd: MI355X-native particle<->grid interpolation (P2G / G2P) with
 * Cabana's operator semantics, behind a plain C ABI.
 *
 * All pointers are device pointers unless stated otherwise.  Every entry
 * point is stream ordered and asynchronous; nothing here allocates or
 * copies user data.  Calls return a status code and never throw.
 *
 * Reference interfaces (paths relative to the Cabana source tree):
 *   grid/src/Cabana_Grid_Interpolation.hpp   p2g / g2p drivers and functors
 *   grid/src/Cabana_Grid_Splines.hpp         Spline<0..6>, evaluateSpline
 *   grid/src/Cabana_Grid_HaloBase.hpp,
 *   grid/src/impl/Cabana_Grid_Halo_Mpi.hpp   node halo gather / scatter(Sum)
 *   core/src/Cabana_LinkedCellList.hpp,
 *   core/src/Cabana_Sort.hpp                 cell binning and permute
 *   core/src/Cabana_Slice.hpp                AoSoA slice addressing
 */
#ifndef CABANA_AMD_H
#define CABANA_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* 2: cabana_amd_binning gained `serial`.
 * 3: serials are issued by cabana_amd_new_binning_serial. */
#define CABANA_AMD_ABI_VERSION 3

typedef struct cabana_amd_context cabana_amd_context;

/* A hipStream_t (0 = the legacy default stream). */
typedef void* cabana_amd_stream;

typedef enum
{
    CABANA_AMD_OK = 0,
    CABANA_AMD_ERROR_INVALID_ARGUMENT = 1,
    CABANA_AMD_ERROR_UNSUPPORTED = 2,
    CABANA_AMD_ERROR_OUT_OF_MEMORY = 3,
    CABANA_AMD_ERROR_DEVICE = 4,
    CABANA_AMD_ERROR_OUT_OF_DOMAIN = 5
} cabana_amd_status;

typedef enum
{
    CABANA_AMD_FLOAT64 = 0,
    CABANA_AMD_FLOAT32 = 1,
    CABANA_AMD_INT32 = 2,
    CABANA_AMD_INT64 = 3
} cabana_amd_dtype;

/* Mesh entity of a grid array (Cabana_Grid_Types.hpp Node, Cell, Face<D>,
 * Edge<D>).  Per dimension an entity sits at the nodes (index i at
 * ghost_low + i * dx; num_cell + 2 * halo + 1 of them) or at the cell
 * centres (ghost_low + (i + 1/2) dx; num_cell + 2 * halo of them): Node at
 * nodes everywhere, Cell at centres everywhere, Face<D> at nodes in D only,
 * Edge<D> at centres in D only (Cabana_Grid_LocalMesh.hpp:176-237). */
typedef enum
{
    CABANA_AMD_NODE = 0,
    CABANA_AMD_CELL = 1,
    CABANA_AMD_FACE_I = 2,
    CABANA_AMD_FACE_J = 3,
    CABANA_AMD_FACE_K = 4,
    CABANA_AMD_EDGE_I = 5,
    CABANA_AMD_EDGE_J = 6,
    CABANA_AMD_EDGE_K = 7
} cabana_amd_entity;

/* Point-to-grid operators: the functors built by createScalarValueP2G,
 * createVectorValueP2G, createScalarGradientP2G, createVectorDivergenceP2G
 * and createTensorDivergenceP2G (Cabana_Grid_Interpolation.hpp:1324-1566). */
typedef enum
{
    CABANA_AMD_SCALAR_VALUE_P2G = 0,
    CABANA_AMD_VECTOR_VALUE_P2G = 1,
    CABANA_AMD_SCALAR_GRADIENT_P2G = 2,
    CABANA_AMD_VECTOR_DIVERGENCE_P2G = 3,
    CABANA_AMD_TENSOR_DIVERGENCE_P2G = 4
} cabana_amd_p2g_op;

/* Grid-to-point operators: createScalarValueG2P, createVectorValueG2P,
 * createScalarGradientG2P, createVectorGradientG2P, createVectorDivergenceG2P
 * (Cabana_Grid_Interpolation.hpp:938-1182). */
typedef enum
{
    CABANA_AMD_SCALAR_VALUE_G2P = 0,
    CABANA_AMD_VECTOR_VALUE_G2P = 1,
    CABANA_AMD_SCALAR_GRADIENT_G2P = 2,
    CABANA_AMD_VECTOR_GRADIENT_G2P = 3,
    CABANA_AMD_VECTOR_DIVERGENCE_G2P = 4
} cabana_amd_g2p_op;

/* One rank's uniform mesh: createUniformGlobalMesh(low, high, cell_size) +
 * createGlobalGrid(periodic) + createLocalGrid(halo_cell_width)
 * (Cabana_Grid_GlobalMesh.hpp:59-102, Cabana_Grid_LocalGrid.hpp:607).  The
 * number of cells per dimension is rint((high - low) / cell_size).  entity
 * (a cabana_amd_entity; 0, Node, when zero-initialised) is the entity of the
 * grid arrays a call passes with it: their extents, owned ranges and halo
 * bands, and the positions the splines weigh (SplineData of the array's
 * entity, Cabana_Grid_Interpolation.hpp:874).  The cell-tiled kernels serve
 * Node arrays; the other entities run the per-particle kernels. */
typedef struct
{
    double low_corner[3];
    double high_corner[3];
    double cell_size[3];
    int32_t halo_cell_width;
    int32_t periodic[3];
    int32_t entity;
} cabana_amd_mesh;

/* Particle field (a Cabana slice or a Kokkos View indexed (p, d0[, d1])).
 * Element (p, c), with c = d0 * D1 + d1, is at
 *     data[(p / vector_length) * soa_stride + (p % vector_length)
 *          + vector_length * c]
 * which is Cabana_Slice.hpp:119-190.  An AoSoA slice has vector_length = VL
 * and soa_stride = slice.stride(0); an (n, ncomp) row-major array has
 * vector_length = 1, soa_stride = ncomp; a column-major one has
 * vector_length = soa_stride = n.  vector_length must be a power of two (as
 * for every Cabana AoSoA) or at least num_particle, num_particle below 2^31
 * and every element index below 2^32 (CABANA_AMD_ERROR_UNSUPPORTED
 * otherwise). */
typedef struct
{
    void* data;
    int64_t num_particle;
    int32_t num_comp;
    int32_t dtype; /* cabana_amd_dtype */
    int64_t vector_length;
    int64_t soa_stride;
} cabana_amd_particle_field;

/* Node array over the ghosted local index space (Array::view(),
 * Cabana_Grid_Array.hpp:227-304): extent = (n_i, n_j, n_k, dofs) with
 * n_d = num_cell_d + 2 * halo_cell_width + 1; strides are in elements, so
 * both LayoutLeft and LayoutRight views are accepted. */
typedef struct
{
    void* data;
    int32_t extent[4];
    int64_t stride[4];
    int32_t dtype; /* cabana_amd_dtype */
} cabana_amd_grid_field;

/* Cell binning of the particles (the output of
 * cabana_amd_linked_cell_build followed by cabana_amd_permute), over the
 * mesh's own cells.  The interpolation calls then walk particles [begin, end)
 * as cell-sorted bricks; the binning only affects speed: particles outside
 * [begin, end), and particles that moved after the binning was built, are
 * interpolated exactly like the rest (through global memory instead of the
 * brick's LDS window).  counts/offsets must still partition [begin, end).
 * NULL means "no ordering is known".
 * serial: 0, or a value returned by cabana_amd_new_binning_serial for this
 * build of counts/offsets.  Calls that pass the same non-zero serial (and
 * the same counts, offsets, begin and end) reuse the brick part tables that
 * an earlier call built from it on this context, in stream order; the owner
 * of the binning takes a new serial whenever counts or offsets change.  A
 * cabana_amd_linked_cell_build into counts/offsets retires every table built
 * from their earlier contents, whatever serial the caller passes later, and a
 * serial the library never issued is rejected (CABANA_AMD_ERROR_INVALID_ARGUMENT).
 * Zero-initialise the struct (= {0} / {}) so that a caller that does not set
 * the field passes 0. */
typedef struct
{
    const int32_t* counts;  /* per cell */
    const int64_t* offsets; /* exclusive scan of counts */
    int64_t begin;
    int64_t end;
    int32_t num_bin[3];     /* must equal the mesh's cells per dimension */
    uint64_t serial;        /* 0: rebuild the part tables in every call */
} cabana_amd_binning;

typedef struct
{
    int32_t op; /* cabana_amd_p2g_op */
    double multiplier;
    cabana_amd_particle_field values; /* read */
    cabana_amd_grid_field grid;       /* accumulated into, then folded */
} cabana_amd_p2g_term;

typedef struct
{
    int32_t op; /* cabana_amd_g2p_op */
    double multiplier;
    cabana_amd_grid_field grid;       /* ghosts filled, then read */
    cabana_amd_particle_field values; /* accumulated into */
} cabana_amd_g2p_term;

/* ----------------------------------------------------------------------- */
/* Context                                                                  */
/* ----------------------------------------------------------------------- */

/* Creates a context bound to a HIP device.  The context owns reusable device
 * scratch (sort keys, tile lists) and per-call diagnostics. */
cabana_amd_status cabana_amd_context_create( int32_t device,
                                             cabana_amd_context** ctx );
cabana_amd_status cabana_amd_context_destroy( cabana_amd_context* ctx );

/* Text of the last error seen by this context (or by the calling thread
 * when ctx is NULL).  Never NULL. */
const char* cabana_amd_last_error( const cabana_amd_context* ctx );

int32_t cabana_amd_abi_version( void );

/* A binning serial no earlier call returned in this process (never 0;
 * thread safe).  No reference counterpart: it keys the library's reuse of the
 * brick part tables it derives from one LinkedCellList build. */
uint64_t cabana_amd_new_binning_serial( void );

/* Waits for the stream, then reports and clears the count of particles that
 * were skipped because their spline stencil left the ghosted grid (Cabana
 * has no such check: such particles are undefined behaviour there).
 * Returns CABANA_AMD_ERROR_OUT_OF_DOMAIN when the count is non-zero. */
cabana_amd_status cabana_amd_check_domain( cabana_amd_context* ctx,
                                           cabana_amd_stream stream,
                                           int64_t* num_skipped );

/* ----------------------------------------------------------------------- */
/* Interpolation                                                            */
/* ----------------------------------------------------------------------- */

/* p2g(functor, points, num_point, Spline<order>, halo, array)
 * (Cabana_Grid_Interpolation.hpp:1217-1313) for 1 to 16 functors.  With
 * cell tiles (bins given, orders 1 to 3) ScalarValue + VectorValue (+
 * TensorDivergence), in any order, run as one fused pass and any other
 * term as a pass of its own; without them every set of terms is one pass
 * over the particles (the reference benchmark's fused_p2g,
 * benchmark/grid/Cabana_Grid_InterpolationPerformance.cpp:25-77).  Each term ADDS its contributions into its grid
 * (owned and ghost nodes), then every distinct grid is folded once: the
 * ghost nodes within halo_width are summed into their periodic owners
 * (Halo::scatter with ScatterReduce::Sum).  Ghost nodes keep their partial
 * sums.  halo_width < 0 means the full halo cell width.  The spline order is
 * 0 to 6 (Spline<0> .. Spline<6>; orders 1 to 3 have the cell-tiled kernels
 * for cell-sorted particles, orders 0 and 4 to 6 run per particle); points
 * are FLOAT64 (num_comp 3).  Precision: FLOAT64 values and
 * grids give the plain FP64 path; FLOAT32 values and grids give the mixed
 * path (FP64 cell-local offsets, FP32 weights and accumulation). */
cabana_amd_status cabana_amd_p2g( cabana_amd_context* ctx,
                                  const cabana_amd_mesh* mesh,
                                  int32_t spline_order,
                                  const cabana_amd_particle_field* points,
                                  int32_t num_terms,
                                  const cabana_amd_p2g_term* terms,
                                  int32_t halo_width,
                                  const cabana_amd_binning* bins,
                                  cabana_amd_stream stream );

/* No reference counterpart: cabana_amd_p2g for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments, semantics and results
 * as cabana_amd_p2g; the call allocates nothing (no hipMalloc/hipFree, no
 * part-table cache, kernel timing off).  The brick part table of a binned
 * (tiled) call is built in a fixed device array of
 * CABANA_AMD_NO_ALLOC_SCRATCH_BYTES per device; a call whose table would
 * not fit returns CABANA_AMD_ERROR_INVALID_ARGUMENT (about 1.3 M bricks plus
 * parts: a 256^3 mesh at any density fits).  Every call on one device uses
 * that device's array: calls that overlap on one device (several streams or
 * host threads) race and must be serialised by the caller.  The context's
 * device must be current. */
#define CABANA_AMD_NO_ALLOC_SCRATCH_BYTES ( (size_t)32 << 20 )
cabana_amd_status cabana_amd_p2g_no_alloc( cabana_amd_context* ctx,
                                           const cabana_amd_mesh* mesh,
                                           int32_t spline_order,
                                           const cabana_amd_particle_field* points,
                                           int32_t num_terms,
                                           const cabana_amd_p2g_term* terms,
                                           int32_t halo_width,
                                           const cabana_amd_binning* bins,
                                           cabana_amd_stream stream );

/* g2p(array, halo, points, num_point, Spline<order>, functor)
 * (Cabana_Grid_Interpolation.hpp:831-927) for 1 to 16 functors.  Every
 * distinct grid first gets its ghosts filled from their periodic owners
 * (Halo::gather), then each term ADDS multiplier * result into its particle
 * field.  With cell tiles VectorValue + VectorGradient from one grid, in
 * either order, run as one fused pass and other terms one pass each;
 * without them every set is one pass (the reference benchmark's fused_g2p,
 * Cabana_Grid_InterpolationPerformance.cpp:81-127). */
cabana_amd_status cabana_amd_g2p( cabana_amd_context* ctx,
                                  const cabana_amd_mesh* mesh,
                                  int32_t spline_order,
                                  const cabana_amd_particle_field* points,
                                  int32_t num_terms,
                                  const cabana_amd_g2p_term* terms,
                                  int32_t halo_width,
                                  const cabana_amd_binning* bins,
                                  cabana_amd_stream stream );

/* g2p2g(scalar_p2g, points, num_point, halo, scalar_array, scalar_g2p)
 * (benchmark/grid/Cabana_Grid_InterpolationPerformance.cpp:130-174) for
 * any terms and spline order: Halo::gather of every distinct G2P grid, then
 * for every particle the G2P terms in order (each adds into its particle
 * field) followed by the P2G terms (each reads its particle field as the G2P
 * terms left it and adds into its grid), then Halo::scatter(Sum) of every
 * distinct P2G grid.  Every G2P read sees the grids as gathered at the
 * start of the call, as the reference's duplicated (Serial, OpenMP)
 * scatter views do; a grid that is both read and added into is read from a
 * device copy.  Without cell tiles (bins NULL, or orders 0 and 4 to 6) this
 * is one pass over the particles; with them, the G2P passes then the P2G
 * passes of cabana_amd_g2p and cabana_amd_p2g. */
cabana_amd_status cabana_amd_g2p2g(
    cabana_amd_context* ctx, const cabana_amd_mesh* mesh, int32_t spline_order,
    const cabana_amd_particle_field* points, int32_t num_g2p_terms,
    const cabana_amd_g2p_term* g2p_terms, int32_t num_p2g_terms,
    const cabana_amd_p2g_term* p2g_terms, int32_t halo_width,
    const cabana_amd_binning* bins, cabana_amd_stream stream );

/* ----------------------------------------------------------------------- */
/* Adjoints (no reference counterpart: Cabana has no differentiation)       */
/* ----------------------------------------------------------------------- */

typedef struct
{
    int32_t op; /* cabana_amd_p2g_op */
    double multiplier;
    cabana_amd_particle_field values;     /* forward input (read) */
    cabana_amd_grid_field grid_bar;       /* cotangent of the folded grid */
    cabana_amd_particle_field values_bar; /* += dL/dvalues; data may be NULL */
} cabana_amd_p2g_adjoint_term;

typedef struct
{
    int32_t op; /* cabana_amd_g2p_op */
    double multiplier;
    cabana_amd_grid_field grid;           /* forward input (ghosts filled) */
    cabana_amd_particle_field values_bar; /* cotangent of the output */
    cabana_amd_grid_field grid_bar;       /* += dL/dgrid; data may be NULL */
} cabana_amd_g2p_adjoint_term;

/* Adjoint of cabana_amd_p2g: a G2P-shaped gather.  Only the owned nodes of
 * each grid_bar are read as the cotangent; its ghost nodes are overwritten
 * with the transpose of the fold (owned value copied into each periodic
 * ghost within halo_width, zero elsewhere).  values_bar += dL/dvalues and,
 * when points_bar is not NULL, points_bar (FLOAT64, 3 comps) += dL/dpoints
 * through the spline weight derivatives.  Weights are recomputed. */
cabana_amd_status cabana_amd_p2g_adjoint(
    cabana_amd_context* ctx, const cabana_amd_mesh* mesh,
    int32_t spline_order, const cabana_amd_particle_field* points,
    int32_t num_terms, const cabana_amd_p2g_adjoint_term* terms,
    int32_t halo_width, const cabana_amd_particle_field* points_bar,
    const cabana_amd_binning* bins, cabana_amd_stream stream );

/* Adjoint of cabana_amd_g2p: a P2G-shaped scatter.  The ghosts of each
 * grid are filled as in the forward call.  Each distinct grid_bar has its
 * ghost nodes zeroed, receives the scattered cotangents, and is folded, so
 * its owned nodes get += dL/dgrid.  points_bar (optional) += dL/dpoints. */
cabana_amd_status cabana_amd_g2p_adjoint(
    cabana_amd_context* ctx, const cabana_amd_mesh* mesh,
    int32_t spline_order, const cabana_amd_particle_field* points,
    int32_t num_terms, const cabana_amd_g2p_adjoint_term* terms,
    int32_t halo_width, const cabana_amd_particle_field* points_bar,
    const cabana_amd_binning* bins, cabana_amd_stream stream );

/* The gather-shaped parts of both adjoints of a transfer pair in one call:
 * cabana_amd_p2g_adjoint of p2g_terms, and the position part of
 * cabana_amd_g2p_adjoint of g2p_terms (their grid_bar is neither read nor
 * written; run cabana_amd_g2p_adjoint with points_bar NULL for the grid
 * part, before or after).  points_bar (required) += both position
 * cotangents; the result equals cabana_amd_p2g_adjoint followed by
 * cabana_amd_g2p_adjoint with grid_bar NULL, up to the order of the
 * additions into points_bar.  For the MPM set (P2G ScalarValue,
 * VectorValue, TensorDivergence; G2P VectorValue, VectorGradient of one
 * grid that is none of the grid_bar) it is one pass over the particles,
 * which reads each position and position cotangent once; other sets run
 * the passes of the two calls. */
cabana_amd_status cabana_amd_adjoint_gather(
    cabana_amd_context* ctx, const cabana_amd_mesh* mesh, int32_t spline_order,
    const cabana_amd_particle_field* points, int32_t num_p2g_terms,
    const cabana_amd_p2g_adjoint_term* p2g_terms, int32_t num_g2p_terms,
    const cabana_amd_g2p_adjoint_term* g2p_terms, int32_t halo_width,
    const cabana_amd_particle_field* points_bar, const cabana_amd_binning* bins,
    cabana_amd_stream stream );

/* ----------------------------------------------------------------------- */
/* Halo (single rank, periodic dimensions exchange with themselves)          */
/* ----------------------------------------------------------------------- */

/* Halo::gather (Cabana_Grid_Halo_Mpi.hpp:76-157): ghost = periodic owner. */
cabana_amd_status cabana_amd_halo_gather( cabana_amd_context* ctx,
                                          const cabana_amd_mesh* mesh,
                                          const cabana_amd_grid_field* grid,
                                          int32_t halo_width,
                                          cabana_amd_stream stream );

/* Halo::scatter(ScatterReduce::Sum) (Cabana_Grid_Halo_Mpi.hpp:166-247):
 * owner += every periodic ghost image of it. */
cabana_amd_status cabana_amd_halo_scatter_sum(
    cabana_amd_context* ctx, const cabana_amd_mesh* mesh,
    const cabana_amd_grid_field* grid, int32_t halo_width,
    cabana_amd_stream stream );

/* ----------------------------------------------------------------------- */
/* Binning                                                                  */
/* ----------------------------------------------------------------------- */

/* Host-side: number of bins per dimension of LinkedCellList's Cartesian grid
 * (CartesianGrid::init, core/src/impl/Cabana_CartesianGrid.hpp:61-73). */
cabana_amd_status cabana_amd_linked_cell_num_bins( const double grid_delta[3],
                                                   const double grid_min[3],
                                                   const double grid_max[3],
                                                   int32_t num_bin[3] );

/* LinkedCellList::build(positions, begin, end)
 * (core/src/Cabana_LinkedCellList.hpp:651-739), reproducing the Serial
 * result bit for bit: counts[c] = particles in cell c, offsets = exclusive
 * scan, permutes[offsets[c] + r] = the r-th particle (ascending index) of
 * cell c, as absolute particle indices.  counts and offsets hold one entry
 * per bin; permutes holds end - begin entries.  Particles outside the
 * Cartesian grid make the call return CABANA_AMD_ERROR_OUT_OF_DOMAIN (the
 * reference indexes out of bounds). */
cabana_amd_status cabana_amd_linked_cell_build(
    cabana_amd_context* ctx, const cabana_amd_particle_field* positions,
    int64_t begin, int64_t end, const double grid_delta[3],
    const double grid_min[3], const double grid_max[3], int32_t* counts,
    int64_t* offsets, int64_t* permutes, cabana_amd_stream stream );

/* permute(binning_data, slice) (core/src/Cabana_Sort.hpp:600-656):
 * field[i] = old field[permutes[i - begin]] for i in [begin, end), bitwise,
 * for every component.  Any dtype. */
cabana_amd_status cabana_amd_permute( cabana_amd_context* ctx,
                                      const int64_t* permutes, int64_t begin,
                                      int64_t end,
                                      const cabana_amd_particle_field* field,
                                      cabana_amd_stream stream );

/* permute(binning_data, aosoa) (core/src/Cabana_Sort.hpp:549-589): every
 * member of a particle set at once, field[i] = old field[permutes[i - begin]]
 * for i in [begin, end) and every member, bitwise.  The members must not
 * share memory (as the members of one AoSoA do not).  Row-contiguous
 * members (vector length 1, 4-byte multiples) go through one fused gather
 * that reads the permutation once; the others are permuted one at a time.
 * Any dtype. */
cabana_amd_status cabana_amd_permute_members(
    cabana_amd_context* ctx, const int64_t* permutes, int64_t begin, int64_t end,
    const cabana_amd_particle_field* fields, int32_t num_fields,
    cabana_amd_stream stream );

/* ----------------------------------------------------------------------- */
/* Grid output                                                              */
/* ----------------------------------------------------------------------- */

/* Experimental::BovWriter::writeTimeStep(prefix, time_step_index, time,
 * array, gather_array) (grid/src/Cabana_Grid_BovWriter.hpp:180-338) on one
 * rank: writes <prefix>_<index, 6 digits>.dat, the owned entities of every
 * dimension (for nodes including the closing node of each dimension) x
 * fastest with the dofs innermost, in the grid's precision, and the VisIt
 * header <prefix>_<index>.bov (TIME, DATA_FILE, DATA_SIZE, DATA_FORMAT,
 * VARIABLE = label, DATA_ENDIAN, CENTERING, BRICK_ORIGIN, BRICK_SIZE,
 * DATA_COMPONENTS).  For a node array with a periodic dimension and
 * gather_array set, a width-0 node halo gather first fills the closing
 * nodes from their periodic owners (a side effect on the array, as there).
 * Waits for the stream. */
cabana_amd_status cabana_amd_bov_write_time_step(
    cabana_amd_context* ctx, const cabana_amd_mesh* mesh,
    const cabana_amd_grid_field* grid, int32_t entity, const char* label,
    const char* prefix, int32_t time_step_index, double time,
    int32_t gather_array, cabana_amd_stream stream );

/* ----------------------------------------------------------------------- */
/* Particle migration (one rank)                                             */
/* ----------------------------------------------------------------------- */

/* particleMigrate(local_grid, positions, particles, min_halo_width,
 * force_migrate) (grid/src/Cabana_Grid_ParticleDistributor.hpp:330-352)
 * on one rank.  Unless force_migrate is set, nothing happens when every
 * particle lies within the band migrateCount checks (:170-221: the ghosted
 * domain moved in by min_halo_width cells); *migrated then stays 0.
 * Otherwise (*migrated = 1) getMigrateDestinations (:87-148) runs: every
 * periodic coordinate above the global high corner moves down by the
 * extent, one below the low corner moves up (one shift), and particles
 * beyond the domain in a non-periodic dimension, whose destination rank is
 * -1, are removed by compacting every member in place, the remaining
 * particles keeping their order (Distributor::migrate, Serial order,
 * core/src/Cabana_Distributor.hpp:80-101).  positions (FLOAT64, num_comp 3)
 * are shifted in place; list them among the members too when they belong
 * to the particle set.  *num_particle returns the number of particles
 * after the call.  Every member holds at least positions->num_particle
 * particles; members must not share memory.  Waits for the stream. */
cabana_amd_status cabana_amd_particle_migrate(
    cabana_amd_context* ctx, const cabana_amd_mesh* mesh,
    const cabana_amd_particle_field* positions,
    const cabana_amd_particle_field* members, int32_t num_members,
    int32_t min_halo_width, int32_t force_migrate, int64_t* num_particle,
    int32_t* migrated, cabana_amd_stream stream );

/* ----------------------------------------------------------------------- */
/* The library's own binning                                                */
/* ----------------------------------------------------------------------- */

/* Calls that pass no binning (bins NULL, the drop-in shape of Cabana's p2g
 * and g2p) with spline orders 1 to 3 bin the particles by owned cell
 * themselves on the call's stream (cell keys, a stable radix sort, counts
 * and offsets, in context scratch) and run the cell-tiled kernels, reading
 * the particles in cell order through the sorted permutation; the
 * particles are not moved.  Results are those of the per-particle path up
 * to the order of floating-point additions.  enable = 0 turns this off for
 * the context (every unbinned call then runs the per-particle kernels with
 * global atomics); it is on by default.  No reference counterpart. */
cabana_amd_status cabana_amd_set_auto_binning( cabana_amd_context* ctx, int32_t enable );

/* LDS atomic-add throughput probe (no reference counterpart: the
 * measurement behind the scatter's FP64 LDS window).  Runs 16 waves per CU,
 * each lane adding iters times into its own LDS word, with ds_add_f64
 * (fp64 != 0) or ds_add_f32, and returns the CU cycles per wave-wide add
 * (kernel time x clock / (16 x iters)).  Synchronous; for measurement only. */
cabana_amd_status cabana_amd_probe_lds_add( cabana_amd_context* ctx, int32_t fp64, int32_t iters,
                                            double* cycles_per_wave_op );

/* Deterministic P2G (enable = 1; off by default).  cabana_amd_p2g and the
 * P2G half of cabana_amd_g2p2g then bin the particles themselves and run a
 * pull pass per term: every node of the grid sums the contributions of the
 * particles of the cells around it, cell by cell in cardinal order and in
 * binned order inside a cell, in FP64, and is written once; the fold is
 * deterministic too.  The same inputs then give bitwise the same grids
 * from call to call (the reference's atomic ScatterView does not).  Node
 * arrays only (CABANA_AMD_ERROR_UNSUPPORTED otherwise); slower than the
 * default path; the adjoint scatters and cabana_amd_p2g_no_alloc are not
 * affected.  No reference counterpart. */
cabana_amd_status cabana_amd_set_deterministic( cabana_amd_context* ctx, int32_t enable );

/* ----------------------------------------------------------------------- */
/* Kernel timing (HIP events around every launch of this context)           */
/* ----------------------------------------------------------------------- */

cabana_amd_status cabana_amd_profile_enable( cabana_amd_context* ctx,
                                             int32_t enable );
cabana_amd_status cabana_amd_profile_reset( cabana_amd_context* ctx );
/* Synchronizes the recorded events, then returns the number of distinct
 * kernel names seen. */
cabana_amd_status cabana_amd_profile_count( cabana_amd_context* ctx,
                                            int32_t* count );
cabana_amd_status cabana_amd_profile_entry( cabana_amd_context* ctx,
                                            int32_t index, char* name,
                                            int32_t name_capacity,
                                            int64_t* launches,
                                            double* total_ms );

#ifdef __cplusplus
}
#endif

#endif
