/*
 * Oracle for the particle<->grid interpolation path: a plain-C, FP64, CPU
 * restatement of Cabana's Serial algorithm.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code.  It is used by tests/ as the checker, by the smoke check
 * of the package entry point, and by bench.py's cpu_baseline leg.
 *
 * Each function cites the reference file:line whose arithmetic it follows
 * (paths relative to the Cabana source tree).
 */
#ifndef CABANA_AMD_ORACLE_H
#define CABANA_AMD_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Single-rank uniform local mesh, Node entity.
 * grid/src/Cabana_Grid_GlobalMesh.hpp:59-80,160-165 (cell count = rint)
 * grid/src/Cabana_Grid_LocalMesh.hpp:61-120 (ghost low corner)
 * grid/src/Cabana_Grid_LocalGrid_impl.hpp:430-467 (node index spaces) */
typedef struct
{
    double global_low[3];
    double global_high[3];
    double cell_size[3];
    int32_t num_cell[3];  /* owned = global cells (one rank) */
    int32_t halo;         /* halo cell width h of the local grid */
    int32_t periodic[3];
    double ghost_low[3];  /* own low corner - h * cell size */
    int32_t num_node[3];  /* ghosted entity extent: num_cell + 2h + 1 - stag */
    int32_t entity;       /* 0 Node, 1 Cell, 2-4 Face<I,J,K>, 5-7 Edge<I,J,K> */
    int32_t stag[3];      /* 1 where the entity sits at cell centres */
} orc_mesh;

/* Particle field: element (p, c) lives at
 *   data[(p / vl) * soa_stride + (p % vl) + vl * c]
 * which is Cabana's AoSoA slice addressing (core/src/Cabana_Slice.hpp:119-190,
 * core/src/impl/Cabana_Index.hpp:62-96) with c = d0 * D1 + d1 flattened. */
typedef struct
{
    double* data;
    int64_t n;
    int32_t ncomp;
    int64_t vl;
    int64_t soa_stride;
} orc_pfield;

/* Ghosted node array (i, j, k, dof) with element strides.
 * grid/src/Cabana_Grid_Array.hpp:227-304 */
typedef struct
{
    double* data;
    int32_t ext[4];
    int64_t stride[4];
} orc_grid;

/* Per-point spline data, grid/src/Cabana_Grid_Splines.hpp:1051-1134. */
/* Knots of the highest spline order (Spline<6>). */
#define ORC_MAX_KNOT 7

typedef struct
{
    int order;
    int num_knot;
    double dx[3];
    double rdx[3];
    double x[3];
    int s[3][ORC_MAX_KNOT];
    double w[3][ORC_MAX_KNOT];
    double g[3][ORC_MAX_KNOT];
    double d[3][ORC_MAX_KNOT];
} orc_spline_data;

/* Operator ids: the ten functors of grid/src/Cabana_Grid_Interpolation.hpp. */
enum
{
    ORC_P2G_SCALAR_VALUE = 0,      /* :1324-1362 */
    ORC_P2G_VECTOR_VALUE = 1,      /* :1373-1413 */
    ORC_P2G_SCALAR_GRADIENT = 2,   /* :1424-1462 */
    ORC_P2G_VECTOR_DIVERGENCE = 3, /* :1473-1513 */
    ORC_P2G_TENSOR_DIVERGENCE = 4  /* :1524-1566 */
};
enum
{
    ORC_G2P_SCALAR_VALUE = 0,     /* :938-977 */
    ORC_G2P_VECTOR_VALUE = 1,     /* :988-1028 */
    ORC_G2P_SCALAR_GRADIENT = 2,  /* :1039-1079 */
    ORC_G2P_VECTOR_GRADIENT = 3,  /* :1090-1132 */
    ORC_G2P_VECTOR_DIVERGENCE = 4 /* :1143-1182 */
};

int orc_mesh_init( orc_mesh* m, const double low[3], const double high[3],
                   const double cell_size[3], int halo, const int periodic[3] );
void orc_node_coordinates( const orc_mesh* m, const int idx[3], double x[3] );

/* 1-D splines, grid/src/Cabana_Grid_Splines.hpp orders 0-3 (:37-452). */
int orc_spline_num_knot( int order );
double orc_spline_map( int order, double xp, double rdx, double low_x );
void orc_spline_stencil( int order, double x0, int* idx );
void orc_spline_value( int order, double x0, double* w );
void orc_spline_gradient( int order, double x0, double rdx, double* g );
/* Second derivative of the 1-D weights in the physical frame.  Not in the
 * reference; derived by differentiating Spline<N>::gradient. */
void orc_spline_hessian( int order, double x0, double rdx, double* h );

void orc_evaluate_spline( const orc_mesh* m, int order, const double p[3],
                          orc_spline_data* sd );

/* Global drivers (Serial order): p2g adds into the grid then folds ghosts into
 * owners; g2p fills ghosts then adds multiplier*result into the particles. */
int orc_p2g( const orc_mesh* m, int order, int op, double multiplier,
             const orc_pfield* points, const orc_pfield* values,
             orc_grid* grid, int halo_width );
int orc_g2p( const orc_mesh* m, int order, int op, double multiplier,
             orc_grid* grid, int halo_width, const orc_pfield* points,
             orc_pfield* values );

/* Same drivers with OpenMP: P2G uses one private grid copy per thread and a
 * final sum (Kokkos ScatterView duplication on OpenMP). */
int orc_p2g_omp( const orc_mesh* m, int order, int op, double multiplier,
                 const orc_pfield* points, const orc_pfield* values,
                 orc_grid* grid, int halo_width, int nthreads );
int orc_g2p_omp( const orc_mesh* m, int order, int op, double multiplier,
                 orc_grid* grid, int halo_width, const orc_pfield* points,
                 orc_pfield* values, int nthreads );

/* Single-rank node halo, 26 neighbour directions, self messages.
 * grid/src/Cabana_Grid_HaloBase.hpp:219-280,334-443,578-738
 * grid/src/impl/Cabana_Grid_Halo_Mpi.hpp:76-247 */
int orc_halo_gather( const orc_mesh* m, orc_grid* grid, int halo_width );
int orc_halo_scatter_sum( const orc_mesh* m, orc_grid* grid, int halo_width );

/* Adjoints (new; no reference counterpart).  Inputs are the forward inputs
 * plus output cotangents; cotangents are ACCUMULATED into the outputs.
 *   p2g adjoint: grid_bar holds the cotangent of the owned grid after the
 *     fold; values_bar += dL/dvalues, points_bar += dL/dpoints.
 *   g2p adjoint: values_bar is the cotangent of the particle output;
 *     grid_bar += dL/dgrid (owned nodes; ghost nodes get partial sums),
 *     points_bar += dL/dpoints.  The grid is read after a ghost fill.
 * Either *_bar pointer may be NULL to skip that output. */
int orc_p2g_adjoint( const orc_mesh* m, int order, int op, double multiplier,
                     const orc_pfield* points, const orc_pfield* values,
                     orc_grid* grid_bar, int halo_width,
                     orc_pfield* values_bar, orc_pfield* points_bar,
                     int nthreads );
int orc_g2p_adjoint( const orc_mesh* m, int order, int op, double multiplier,
                     orc_grid* grid, int halo_width, const orc_pfield* points,
                     const orc_pfield* values_bar, orc_grid* grid_bar,
                     orc_pfield* points_bar, int nthreads );

/* Cartesian binning grid, core/src/impl/Cabana_CartesianGrid.hpp:61-73,
 * 171-182, 226-231, 279-283. */
typedef struct
{
    double min[3], max[3], dx[3], rdx[3];
    int32_t nx[3];
} orc_cart_grid;

void orc_cart_grid_init( orc_cart_grid* g, const double min[3],
                         const double max[3], const double delta[3] );
void orc_cart_locate( const orc_cart_grid* g, const double p[3], int ijk[3] );
int64_t orc_cart_cardinal( const orc_cart_grid* g, const int ijk[3] );

/* LinkedCellList::build in Serial order, core/src/Cabana_LinkedCellList.hpp:651-739.
 * counts/offsets: one per bin; permutes: end - begin entries (absolute ids). */
int orc_lcl_build( const orc_pfield* positions, int64_t begin, int64_t end,
                   const double delta[3], const double min[3],
                   const double max[3], int32_t* counts, int64_t* offsets,
                   int64_t* permutes );
/* permute(binning data, slice), core/src/Cabana_Sort.hpp:600-656 */
int orc_permute( const int64_t* permutes, int64_t begin, int64_t end,
                 orc_pfield* field );

/* Makes the mesh's arrays those of another entity (the index spaces and
 * coordinates of grid/src/Cabana_Grid_LocalGrid_impl.hpp and
 * Cabana_Grid_LocalMesh.hpp:176-237).  Returns 0 on success. */
int orc_mesh_set_entity( orc_mesh* m, int entity );

/* particleMigrate on one rank over row-major positions x[n][3]: -1 when
 * nothing migrates, else the kept count with keep[p] set (0 = removed). */
int64_t orc_particle_migrate( const orc_mesh* m, double* x, int64_t n, int min_halo_width,
                              int force, int32_t* keep );

/* BovWriter::writeTimeStep on one rank (FP64 grids); entity 0 = Node,
 * 1 = Cell.  Returns 0 on success. */
int orc_bov_write( const orc_mesh* m, orc_grid* grid, int entity, const char* label,
                   const char* prefix, int time_step_index, double time, int gather_array );

#ifdef __cplusplus
}
#endif

#endif
