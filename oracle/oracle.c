/*
 * Oracle: plain-C FP64 restatement of Cabana's Serial particle<->grid path.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, the package smoke check and the
 * cpu_baseline leg of bench.py.  The product library never links or calls it.
 *
 * Build with -ffp-contract=off so every expression rounds exactly as written.
 * Citations are file:line in the Cabana source tree.
 */
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* Mesh                                                                     */
/* ------------------------------------------------------------------------ */

/* GlobalMesh cell-size constructor, grid/src/Cabana_Grid_GlobalMesh.hpp:83-102:
 * globalNumCell = rint(extent / cell_size) (:160-165); extents must divide
 * within 100 eps.  LocalMesh ghost low corner = own low - h * cell size,
 * grid/src/Cabana_Grid_LocalMesh.hpp:100-104 (own low = global low on one
 * rank, :89-92). */
int orc_mesh_init( orc_mesh* m, const double low[3], const double high[3],
                   const double cell_size[3], int halo, const int periodic[3] )
{
    if ( halo < 0 )
        return 1;
    for ( int d = 0; d < 3; ++d )
    {
        m->global_low[d] = low[d];
        m->global_high[d] = high[d];
        m->cell_size[d] = cell_size[d];
        double extent = high[d] - low[d];
        m->num_cell[d] = (int32_t)rint( extent / cell_size[d] );
        double ext = m->num_cell[d] * cell_size[d];
        if ( fabs( ext - extent ) > 100.0 * 2.220446049250313e-16 )
            return 2;
        if ( m->num_cell[d] < 1 )
            return 3;
        m->periodic[d] = periodic[d] ? 1 : 0;
        double own_low = low[d] + cell_size[d] * 0;
        m->ghost_low[d] = own_low - halo * cell_size[d];
        m->num_node[d] = m->num_cell[d] + 2 * halo + 1;
        m->stag[d] = 0;
    }
    m->halo = halo;
    m->entity = 0;
    return 0;
}

/* Entity of a mesh's arrays, one dimension at a time: Cell sits at cell
 * centres in every dimension, Face<D> at nodes in D and at cell centres in
 * the others, Edge<D> the other way round (LocalMesh::coordinates,
 * Cabana_Grid_LocalMesh.hpp:176-237).  A cell-centred dimension has one
 * entity fewer (faceIndexSpace / edge and cell index spaces,
 * Cabana_Grid_LocalGrid_impl.hpp:239-280, 534-560). */
int orc_mesh_set_entity( orc_mesh* m, int entity )
{
    if ( entity < 0 || entity > 7 )
        return 1;
    for ( int d = 0; d < 3; ++d )
    {
        int s;
        if ( entity == 0 )
            s = 0;
        else if ( entity == 1 )
            s = 1;
        else if ( entity <= 4 )
            s = ( d == entity - 2 ) ? 0 : 1;
        else
            s = ( d == entity - 5 ) ? 1 : 0;
        m->stag[d] = s;
        m->num_node[d] = m->num_cell[d] + 2 * m->halo + 1 - s;
    }
    m->entity = entity;
    return 0;
}

/* LocalMesh::coordinates(Entity, index), grid/src/Cabana_Grid_LocalMesh.hpp
 * :176-237: ghost low + index * dx at nodes, ghost low + (index + 0.5) * dx
 * at cell centres. */
void orc_node_coordinates( const orc_mesh* m, const int idx[3], double x[3] )
{
    for ( int d = 0; d < 3; ++d )
        x[d] = m->stag[d] ? m->ghost_low[d] + ( (double)idx[d] + 0.5 ) * m->cell_size[d]
                          : m->ghost_low[d] + (double)idx[d] * m->cell_size[d];
}

/* ------------------------------------------------------------------------ */
/* 1-D splines                                                              */
/* ------------------------------------------------------------------------ */

int orc_spline_num_knot( int order ) { return order + 1; }

/* Spline<4>, <5>, <6> (Cabana_Grid_Splines.hpp:456-598, :600-761, :763-939)
 * write every knot's value and gradient as a polynomial in
 * xn = x0 - int(x0) - 0.5 with integer coefficients over a per-knot
 * denominator.  These are those polynomials, lowest power first: value
 * numerators and denominators (Spline<4>::value :522-557, Spline<5>::value
 * :668-708, Spline<6>::value :833-881) and gradient ones, in units of rdx
 * (Spline<4>::gradient :568-597, Spline<5>::gradient :720-760,
 * Spline<6>::gradient :893-938). */
static const double hi_val4[5][5] = { { 1, -8, 24, -32, 16 },
                                      { 19, -44, 24, 16, -16 },
                                      { 115, 0, -120, 0, 48 },
                                      { 19, 44, 24, -16, -16 },
                                      { 1, 8, 24, 32, 16 } };
static const double hi_vden4[5] = { 384, 96, 192, 96, 384 };
static const double hi_grad4[5][4] = { { -1, 6, -12, 8 },
                                       { -11, 12, 12, -16 },
                                       { 0, -5, 0, 4 },
                                       { 11, 12, -12, -16 },
                                       { 1, 6, 12, 8 } };
static const double hi_gden4[5] = { 48, 24, 4, 24, 48 };

static const double hi_val5[6][6] = { { 1, -10, 40, -80, 80, -32 },
                                      { 237, -750, 840, -240, -240, 160 },
                                      { 841, -770, -440, 560, 80, -160 },
                                      { 841, 770, -440, -560, 80, 160 },
                                      { 237, 750, 840, 240, -240, -160 },
                                      { 1, 10, 40, 80, 80, 32 } };
static const double hi_vden5[6] = { 3840, 3840, 1920, 1920, 3840, 3840 };
static const double hi_grad5[6][5] = { { -1, 8, -24, 32, -16 },
                                       { -75, 168, -72, -96, 80 },
                                       { -77, -88, 168, 32, -80 },
                                       { 77, -88, -168, 32, 80 },
                                       { 75, 168, 72, -96, -80 },
                                       { 1, 8, 24, 32, 16 } };
static const double hi_gden5[6] = { 384, 384, 192, 192, 384, 384 };

static const double hi_val6[7][7] = {
    { 1, -12, 60, -160, 240, -192, 64 },
    { 361, -1416, 2220, -1600, 240, 384, -192 },
    { 10543, -17340, 4740, 6880, -4080, -960, 960 },
    { 5887, 0, -4620, 0, 1680, 0, -320 },
    { 10543, 17340, 4740, -6880, -4080, 960, 960 },
    { 361, 1416, 2220, 1600, 240, -384, -192 },
    { 1, 12, 60, 160, 240, 192, 64 } };
static const double hi_vden6[7] = { 46080, 23040, 46080, 11520, 46080, 23040, 46080 };
static const double hi_grad6[7][6] = { { -1, 10, -40, 80, -80, 32 },
                                       { -59, 185, -200, 40, 80, -48 },
                                       { -289, 158, 344, -272, -80, 96 },
                                       { 0, -77, 0, 56, 0, -16 },
                                       { 289, 158, -344, -272, 80, 96 },
                                       { 59, 185, 200, 40, -80, -48 },
                                       { 1, 10, 40, 80, 80, 32 } };
static const double hi_gden6[7] = { 3840, 960, 768, 96, 768, 960, 3840 };

/* Numerator table row j (num_coef coefficients, lowest power first) of the
 * value (deriv 0), gradient (1) or second derivative (2, the derivative of
 * the gradient polynomial) of a degree 4-6 spline, evaluated at xn by
 * Horner's rule and divided by the knot's denominator. */
static double hi_poly( int order, int deriv, int j, double xn )
{
    const double* c;
    double den;
    int nc;
    int d = deriv;
    if ( d == 0 )
    {
        nc = order + 1;
        c = order == 4 ? hi_val4[j] : order == 5 ? hi_val5[j] : hi_val6[j];
        den = order == 4 ? hi_vden4[j] : order == 5 ? hi_vden5[j] : hi_vden6[j];
    }
    else
    {
        nc = order;
        c = order == 4 ? hi_grad4[j] : order == 5 ? hi_grad5[j] : hi_grad6[j];
        den = order == 4 ? hi_gden4[j] : order == 5 ? hi_gden5[j] : hi_gden6[j];
        d -= 1;
    }
    /* d = 0: the polynomial itself; d = 1: its derivative in xn. */
    double acc = 0.0;
    for ( int n = nc - 1; n >= d; --n )
        acc = acc * xn + ( d ? n * c[n] : c[n] );
    return acc / den;
}

/* mapToLogicalGrid: dual grid (+0.5) for even orders, primal for odd.
 * Spline<0> :59-63, Spline<1> :141-145, Spline<2> :239-244, Spline<3> :349-353,
 * Spline<4> :477-481, Spline<5> :621-625, Spline<6> :784-788 */
double orc_spline_map( int order, double xp, double rdx, double low_x )
{
    if ( order % 2 == 0 )
        return ( xp - low_x ) * rdx + 0.5;
    return ( xp - low_x ) * rdx;
}

/* stencil: Spline<0> :79-84, <1> :163-169, <2> :264-271, <3> :383-391,
 * <4> :504-512 (int(x0) - 2), <5> :649-659 (int(x0) - 2), <6> :813-824
 * (int(x0) - 3) */
void orc_spline_stencil( int order, double x0, int* idx )
{
    int base = (int)x0;
    if ( order >= 2 )
        base -= order / 2;
    for ( int n = 0; n <= order; ++n )
        idx[n] = base + n;
}

/* value: Spline<0> :95-100, <1> :178-187, <2> :280-298, <3> :400-427,
 * <4>..<6> in hi_poly */
void orc_spline_value( int order, double x0, double* w )
{
    if ( order >= 4 )
    {
        const double xn = x0 - (int)x0 - 0.5;
        for ( int j = 0; j <= order; ++j )
            w[j] = hi_poly( order, 0, j, xn );
    }
    else if ( order == 0 )
    {
        w[0] = 1.0;
    }
    else if ( order == 1 )
    {
        double xn = x0 - (int)x0;
        w[0] = 1.0 - xn;
        w[1] = xn;
    }
    else if ( order == 2 )
    {
        double nine_eights = 9.0 / 8.0;
        double xn = x0 - (int)x0 + 0.5;
        w[0] = 0.5 * xn * xn - 1.5 * xn + nine_eights;
        xn -= 1.0;
        w[1] = -xn * xn + 0.75;
        xn -= 1.0;
        w[2] = 0.5 * xn * xn + 1.5 * xn + nine_eights;
    }
    else
    {
        double one_sixth = 1.0 / 6.0;
        double two_thirds = one_sixth * 4.0;
        double four_thirds = two_thirds * 2.0;
        double xn = x0 - (int)x0 + 1.0;
        double xn2 = xn * xn;
        w[0] = -xn * xn2 * one_sixth + xn2 - 2.0 * xn + four_thirds;
        xn -= 1.0;
        xn2 = xn * xn;
        w[1] = 0.5 * xn * xn2 - xn2 + two_thirds;
        xn -= 1.0;
        xn2 = xn * xn;
        w[2] = -0.5 * xn * xn2 - xn2 + two_thirds;
        xn -= 1.0;
        xn2 = xn * xn;
        w[3] = xn * xn2 * one_sixth + xn2 + 2.0 * xn + four_thirds;
    }
}

/* gradient: Spline<0> :111-117, <1> :199-213, <2> :309-324, <3> :437-451,
 * <4>..<6> in hi_poly */
void orc_spline_gradient( int order, double x0, double rdx, double* g )
{
    if ( order >= 4 )
    {
        const double xn = x0 - (int)x0 - 0.5;
        for ( int j = 0; j <= order; ++j )
            g[j] = hi_poly( order, 1, j, xn ) * rdx;
    }
    else if ( order == 0 )
    {
        g[0] = 0.0;
    }
    else if ( order == 1 )
    {
        g[0] = -rdx;
        g[1] = rdx;
    }
    else if ( order == 2 )
    {
        double xn = x0 - (int)x0 + 0.5;
        g[0] = ( xn - 1.5 ) * rdx;
        xn -= 1.0;
        g[1] = ( -2.0 * xn ) * rdx;
        xn -= 1.0;
        g[2] = ( xn + 1.5 ) * rdx;
    }
    else
    {
        double xn = x0 - (int)x0 + 1.0;
        g[0] = ( -0.5 * xn * xn + 2.0 * xn - 2.0 ) * rdx;
        xn -= 1.0;
        g[1] = ( 1.5 * xn * xn - 2.0 * xn ) * rdx;
        xn -= 1.0;
        g[2] = ( -1.5 * xn * xn - 2.0 * xn ) * rdx;
        xn -= 1.0;
        g[3] = ( 0.5 * xn * xn + 2.0 * xn + 2.0 ) * rdx;
    }
}

/* d/dx of the gradients above (each is a polynomial in xn with dxn/dx = rdx). */
void orc_spline_hessian( int order, double x0, double rdx, double* h )
{
    double r2 = rdx * rdx;
    if ( order >= 4 )
    {
        const double xn = x0 - (int)x0 - 0.5;
        for ( int j = 0; j <= order; ++j )
            h[j] = hi_poly( order, 2, j, xn ) * r2;
    }
    else if ( order <= 1 )
    {
        for ( int n = 0; n <= order; ++n )
            h[n] = 0.0;
    }
    else if ( order == 2 )
    {
        h[0] = r2;
        h[1] = -2.0 * r2;
        h[2] = r2;
    }
    else
    {
        double xn = x0 - (int)x0 + 1.0;
        h[0] = ( -xn + 2.0 ) * r2;
        xn -= 1.0;
        h[1] = ( 3.0 * xn - 2.0 ) * r2;
        xn -= 1.0;
        h[2] = ( -3.0 * xn - 2.0 ) * r2;
        xn -= 1.0;
        h[3] = ( xn + 2.0 ) * r2;
    }
}

/* evaluateSpline, grid/src/Cabana_Grid_Splines.hpp:1283-1342: the low corner
 * is the coordinate of local node 0 and dx the DIFFERENCE of the coordinates
 * of nodes 1 and 0 (:1304-1318), not the stored cell size. */
void orc_evaluate_spline( const orc_mesh* m, int order, const double p[3],
                          orc_spline_data* sd )
{
    int low_id[3] = { 0, 0, 0 };
    int low_id_p1[3] = { 1, 1, 1 };
    double low_x[3], low_x_p1[3];
    orc_node_coordinates( m, low_id, low_x );
    orc_node_coordinates( m, low_id_p1, low_x_p1 );
    sd->order = order;
    sd->num_knot = order + 1;
    for ( int d = 0; d < 3; ++d )
    {
        sd->dx[d] = low_x_p1[d] - low_x[d];
        sd->rdx[d] = 1.0 / sd->dx[d];
        sd->x[d] = orc_spline_map( order, p[d], sd->rdx[d], low_x[d] );
        orc_spline_stencil( order, sd->x[d], sd->s[d] );
        orc_spline_value( order, sd->x[d], sd->w[d] );
        orc_spline_gradient( order, sd->x[d], sd->rdx[d], sd->g[d] );
    }
    /* setSplineData(SplinePhysicalDistance) :1244-1266 */
    for ( int d = 0; d < 3; ++d )
    {
        double offset = low_x[d] - p[d];
        for ( int n = 0; n < sd->num_knot; ++n )
            sd->d[d][n] = offset + sd->s[d][n] * sd->dx[d];
    }
}

/* ------------------------------------------------------------------------ */
/* Field access                                                             */
/* ------------------------------------------------------------------------ */

static inline int64_t pf_off( const orc_pfield* f, int64_t p, int c )
{
    return ( p / f->vl ) * f->soa_stride + ( p % f->vl ) + f->vl * (int64_t)c;
}
static inline double pf_get( const orc_pfield* f, int64_t p, int c )
{
    return f->data[pf_off( f, p, c )];
}
static inline double* pf_ptr( orc_pfield* f, int64_t p, int c )
{
    return f->data + pf_off( f, p, c );
}
static inline double* gd_ptr( const orc_grid* g, int i, int j, int k, int d )
{
    return g->data + i * g->stride[0] + j * g->stride[1] + k * g->stride[2] +
           d * g->stride[3];
}

static int p2g_grid_dofs( int op )
{
    switch ( op )
    {
    case ORC_P2G_SCALAR_VALUE:
    case ORC_P2G_VECTOR_DIVERGENCE:
        return 1;
    default:
        return 3;
    }
}
static int p2g_point_comps( int op )
{
    switch ( op )
    {
    case ORC_P2G_SCALAR_VALUE:
    case ORC_P2G_SCALAR_GRADIENT:
        return 1;
    case ORC_P2G_TENSOR_DIVERGENCE:
        return 9;
    default:
        return 3;
    }
}
static int g2p_grid_dofs( int op )
{
    switch ( op )
    {
    case ORC_G2P_SCALAR_VALUE:
    case ORC_G2P_SCALAR_GRADIENT:
        return 1;
    default:
        return 3;
    }
}
static int g2p_point_comps( int op )
{
    switch ( op )
    {
    case ORC_G2P_SCALAR_VALUE:
    case ORC_G2P_VECTOR_DIVERGENCE:
        return 1;
    case ORC_G2P_VECTOR_GRADIENT:
        return 9;
    default:
        return 3;
    }
}

static int check_grid( const orc_mesh* m, const orc_grid* g, int dofs )
{
    for ( int d = 0; d < 3; ++d )
        if ( g->ext[d] != m->num_node[d] )
            return 1;
    if ( g->ext[3] != dofs )
        return 1;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Point kernels                                                            */
/* ------------------------------------------------------------------------ */

/* Functor prologues scale the point data by the multiplier
 * (Interpolation.hpp:1344-1351, 1394-1401, 1444-1451, 1494-1501, 1545-1555),
 * then the P2G point kernels scatter (:437-458, 495-518, 557-588, 631-665,
 * 712-747). */
static void p2g_point( int op, double multiplier, const orc_pfield* values,
                       int64_t p, const orc_spline_data* sd, orc_grid* grid )
{
    const int K = sd->num_knot;
    const int I = 0, J = 1, Kd = 2;
    if ( op == ORC_P2G_SCALAR_VALUE )
    {
        double pd = pf_get( values, p, 0 ) * multiplier;
        for ( int i = 0; i < K; ++i )
            for ( int j = 0; j < K; ++j )
                for ( int k = 0; k < K; ++k )
                    *gd_ptr( grid, sd->s[I][i], sd->s[J][j], sd->s[Kd][k],
                             0 ) += pd * sd->w[I][i] * sd->w[J][j] *
                                    sd->w[Kd][k];
    }
    else if ( op == ORC_P2G_VECTOR_VALUE )
    {
        double pd[3];
        for ( int d = 0; d < 3; ++d )
            pd[d] = multiplier * pf_get( values, p, d );
        for ( int i = 0; i < K; ++i )
            for ( int j = 0; j < K; ++j )
                for ( int k = 0; k < K; ++k )
                    for ( int d = 0; d < 3; ++d )
                        *gd_ptr( grid, sd->s[I][i], sd->s[J][j], sd->s[Kd][k],
                                 d ) += pd[d] * sd->w[I][i] * sd->w[J][j] *
                                        sd->w[Kd][k];
    }
    else if ( op == ORC_P2G_SCALAR_GRADIENT )
    {
        double pd = pf_get( values, p, 0 ) * multiplier;
        for ( int i = 0; i < K; ++i )
            for ( int j = 0; j < K; ++j )
                for ( int k = 0; k < K; ++k )
                {
                    int a = sd->s[I][i], b = sd->s[J][j], c = sd->s[Kd][k];
                    *gd_ptr( grid, a, b, c, I ) +=
                        pd * sd->g[I][i] * sd->w[J][j] * sd->w[Kd][k];
                    *gd_ptr( grid, a, b, c, J ) +=
                        pd * sd->w[I][i] * sd->g[J][j] * sd->w[Kd][k];
                    *gd_ptr( grid, a, b, c, Kd ) +=
                        pd * sd->w[I][i] * sd->w[J][j] * sd->g[Kd][k];
                }
    }
    else if ( op == ORC_P2G_VECTOR_DIVERGENCE )
    {
        double pd[3];
        for ( int d = 0; d < 3; ++d )
            pd[d] = multiplier * pf_get( values, p, d );
        for ( int i = 0; i < K; ++i )
            for ( int j = 0; j < K; ++j )
                for ( int k = 0; k < K; ++k )
                {
                    double result =
                        pd[I] * sd->g[I][i] * sd->w[J][j] * sd->w[Kd][k] +
                        pd[J] * sd->w[I][i] * sd->g[J][j] * sd->w[Kd][k] +
                        pd[Kd] * sd->w[I][i] * sd->w[J][j] * sd->g[Kd][k];
                    *gd_ptr( grid, sd->s[I][i], sd->s[J][j], sd->s[Kd][k],
                             0 ) += result;
                }
    }
    else
    {
        double pd[3][3];
        for ( int d0 = 0; d0 < 3; ++d0 )
            for ( int d1 = 0; d1 < 3; ++d1 )
                pd[d0][d1] = multiplier * pf_get( values, p, 3 * d0 + d1 );
        for ( int i = 0; i < K; ++i )
            for ( int j = 0; j < K; ++j )
                for ( int k = 0; k < K; ++k )
                {
                    double rg[3] = {
                        sd->g[I][i] * sd->w[J][j] * sd->w[Kd][k],
                        sd->w[I][i] * sd->g[J][j] * sd->w[Kd][k],
                        sd->w[I][i] * sd->w[J][j] * sd->g[Kd][k] };
                    for ( int d1 = 0; d1 < 3; ++d1 )
                        for ( int d0 = 0; d0 < 3; ++d0 )
                            *gd_ptr( grid, sd->s[I][i], sd->s[J][j],
                                     sd->s[Kd][k], d0 ) +=
                                rg[d1] * pd[d0][d1];
                }
    }
}

/* G2P point kernels (:53-72, 108-130, 168-203, 247-282, 333-364) and the
 * accumulate-into-particle functors (:965, 1016, 1067, 1120, 1170). */
static void g2p_point( int op, double multiplier, const orc_grid* grid,
                       int64_t p, const orc_spline_data* sd,
                       orc_pfield* values )
{
    const int K = sd->num_knot;
    const int I = 0, J = 1, Kd = 2;
    if ( op == ORC_G2P_SCALAR_VALUE )
    {
        double result = 0.0;
        for ( int i = 0; i < K; ++i )
            for ( int j = 0; j < K; ++j )
                for ( int k = 0; k < K; ++k )
                    result += *gd_ptr( grid, sd->s[I][i], sd->s[J][j],
                                       sd->s[Kd][k], 0 ) *
                              sd->w[I][i] * sd->w[J][j] * sd->w[Kd][k];
        *pf_ptr( values, p, 0 ) += multiplier * result;
    }
    else if ( op == ORC_G2P_VECTOR_VALUE )
    {
        double result[3] = { 0.0, 0.0, 0.0 };
        for ( int i = 0; i < K; ++i )
            for ( int j = 0; j < K; ++j )
                for ( int k = 0; k < K; ++k )
                    for ( int d = 0; d < 3; ++d )
                        result[d] += *gd_ptr( grid, sd->s[I][i], sd->s[J][j],
                                              sd->s[Kd][k], d ) *
                                     sd->w[I][i] * sd->w[J][j] * sd->w[Kd][k];
        for ( int d = 0; d < 3; ++d )
            *pf_ptr( values, p, d ) += multiplier * result[d];
    }
    else if ( op == ORC_G2P_SCALAR_GRADIENT )
    {
        double result[3] = { 0.0, 0.0, 0.0 };
        for ( int i = 0; i < K; ++i )
            for ( int j = 0; j < K; ++j )
                for ( int k = 0; k < K; ++k )
                {
                    double v = *gd_ptr( grid, sd->s[I][i], sd->s[J][j],
                                        sd->s[Kd][k], 0 );
                    result[I] += v * sd->g[I][i] * sd->w[J][j] * sd->w[Kd][k];
                    result[J] += v * sd->w[I][i] * sd->g[J][j] * sd->w[Kd][k];
                    result[Kd] += v * sd->w[I][i] * sd->w[J][j] * sd->g[Kd][k];
                }
        for ( int d = 0; d < 3; ++d )
            *pf_ptr( values, p, d ) += multiplier * result[d];
    }
    else if ( op == ORC_G2P_VECTOR_GRADIENT )
    {
        double result[3][3];
        memset( result, 0, sizeof( result ) );
        for ( int i = 0; i < K; ++i )
            for ( int j = 0; j < K; ++j )
                for ( int k = 0; k < K; ++k )
                {
                    double rg[3] = {
                        sd->g[I][i] * sd->w[J][j] * sd->w[Kd][k],
                        sd->w[I][i] * sd->g[J][j] * sd->w[Kd][k],
                        sd->w[I][i] * sd->w[J][j] * sd->g[Kd][k] };
                    for ( int d0 = 0; d0 < 3; ++d0 )
                    {
                        double mg = rg[d0];
                        for ( int d1 = 0; d1 < 3; ++d1 )
                            result[d0][d1] +=
                                mg * *gd_ptr( grid, sd->s[I][i], sd->s[J][j],
                                              sd->s[Kd][k], d1 );
                    }
                }
        for ( int d0 = 0; d0 < 3; ++d0 )
            for ( int d1 = 0; d1 < 3; ++d1 )
                *pf_ptr( values, p, 3 * d0 + d1 ) +=
                    multiplier * result[d0][d1];
    }
    else
    {
        double result = 0.0;
        for ( int i = 0; i < K; ++i )
            for ( int j = 0; j < K; ++j )
                for ( int k = 0; k < K; ++k )
                {
                    int a = sd->s[I][i], b = sd->s[J][j], c = sd->s[Kd][k];
                    result += *gd_ptr( grid, a, b, c, I ) * sd->g[I][i] *
                                  sd->w[J][j] * sd->w[Kd][k] +
                              *gd_ptr( grid, a, b, c, J ) * sd->w[I][i] *
                                  sd->g[J][j] * sd->w[Kd][k] +
                              *gd_ptr( grid, a, b, c, Kd ) * sd->w[I][i] *
                                  sd->w[J][j] * sd->g[Kd][k];
                }
        *pf_ptr( values, p, 0 ) += result * multiplier;
    }
}

/* ------------------------------------------------------------------------ */
/* Halo                                                                     */
/* ------------------------------------------------------------------------ */

/* Owned node space bounds on one rank, LocalGrid_impl.hpp:430-450. */
static void owned_bounds( const orc_mesh* m, int d, int* lo, int* hi )
{
    *lo = m->halo;
    *hi = ( m->periodic[d] || m->stag[d] ) ? m->halo + m->num_cell[d]
                                           : m->halo + m->num_cell[d] + 1;
}

/* Shared index spaces, LocalGrid_impl.hpp:470-509 with getBound :283-320:
 *   Own:   min = getBound(owned, -1, off, 0, -w), max = getBound(owned, 1, off, w+1, 0)
 *   Ghost: min = getBound(owned, -1, off, -w, 0), max = getBound(owned, 1, off, 0, w+1)
 * for node-centred dimensions; cell-centred ones have w in place of w+1
 * (:370-398; faces and edges mix the two per dimension). */
static void shared_box( const orc_mesh* m, int ghost, const int off[3], int w,
                        int lo[3], int hi[3] )
{
    for ( int d = 0; d < 3; ++d )
    {
        int omin, omax;
        owned_bounds( m, d, &omin, &omax );
        if ( off[d] == -1 )
        {
            lo[d] = omin + ( ghost ? -w : 0 );
            hi[d] = omin + ( ghost ? 0 : w + 1 - m->stag[d] );
        }
        else if ( off[d] == 0 )
        {
            lo[d] = omin;
            hi[d] = omax;
        }
        else
        {
            lo[d] = omax + ( ghost ? 0 : -w );
            hi[d] = omax + ( ghost ? w + 1 - m->stag[d] : 0 );
        }
    }
}

/* A neighbour exists on one rank iff every non-zero offset is in a periodic
 * dimension (LocalGrid_impl.hpp:65-73 -> GlobalGrid blockRank). */
static int neighbor_valid( const orc_mesh* m, const int off[3] )
{
    for ( int d = 0; d < 3; ++d )
        if ( off[d] != 0 && !m->periodic[d] )
            return 0;
    return 1;
}

static int halo_exchange( const orc_mesh* m, orc_grid* grid, int w, int sum )
{
    if ( w < 0 )
        w = m->halo;
    if ( w > m->halo )
        return 1;
    for ( int d = 0; d < 3; ++d )
        if ( grid->ext[d] != m->num_node[d] )
            return 1;
    const int dofs = grid->ext[3];
    /* Pack every message first (gather packs owned, scatter packs ghosts),
     * then unpack: Halo_Mpi.hpp:90-157 and :176-247. */
    double* bufs[27];
    int los[27][3], his[27][3], dlos[27][3];
    int64_t sizes[27];
    int nmsg = 0;
    for ( int oi = -1; oi <= 1; ++oi )
        for ( int oj = -1; oj <= 1; ++oj )
            for ( int ok = -1; ok <= 1; ++ok )
            {
                int off[3] = { oi, oj, ok };
                if ( oi == 0 && oj == 0 && ok == 0 )
                    continue;
                if ( !neighbor_valid( m, off ) )
                    continue;
                int flip[3] = { -oi, -oj, -ok };
                int slo[3], shi[3], dlo[3], dhi[3];
                /* Message for direction n travels from the sender's space
                 * for n to the receiver's space for flip(n) (tags
                 * HaloBase.hpp:257-262); on one rank both are ours. */
                if ( sum )
                {
                    shared_box( m, 1, off, w, slo, shi );
                    shared_box( m, 0, flip, w, dlo, dhi );
                }
                else
                {
                    shared_box( m, 0, off, w, slo, shi );
                    shared_box( m, 1, flip, w, dlo, dhi );
                }
                int64_t n = (int64_t)dofs;
                for ( int d = 0; d < 3; ++d )
                    n *= ( shi[d] - slo[d] ) > 0 ? ( shi[d] - slo[d] ) : 0;
                double* b = (double*)malloc( sizeof( double ) * ( n ? n : 1 ) );
                int64_t c = 0;
                for ( int i = slo[0]; i < shi[0]; ++i )
                    for ( int j = slo[1]; j < shi[1]; ++j )
                        for ( int k = slo[2]; k < shi[2]; ++k )
                            for ( int l = 0; l < dofs; ++l )
                                b[c++] = *gd_ptr( grid, i, j, k, l );
                bufs[nmsg] = b;
                sizes[nmsg] = n;
                for ( int d = 0; d < 3; ++d )
                {
                    los[nmsg][d] = slo[d];
                    his[nmsg][d] = shi[d];
                    dlos[nmsg][d] = dlo[d];
                    (void)dhi;
                }
                ++nmsg;
            }
    for ( int q = 0; q < nmsg; ++q )
    {
        int64_t c = 0;
        int ni = his[q][0] - los[q][0], nj = his[q][1] - los[q][1],
            nk = his[q][2] - los[q][2];
        for ( int i = 0; i < ni; ++i )
            for ( int j = 0; j < nj; ++j )
                for ( int k = 0; k < nk; ++k )
                    for ( int l = 0; l < dofs; ++l )
                    {
                        double* dst = gd_ptr( grid, dlos[q][0] + i,
                                              dlos[q][1] + j, dlos[q][2] + k,
                                              l );
                        /* unpackOp Sum / Replace, HaloBase.hpp:600-620 */
                        if ( sum )
                            *dst += bufs[q][c++];
                        else
                            *dst = bufs[q][c++];
                    }
        (void)sizes;
        free( bufs[q] );
    }
    return 0;
}

int orc_halo_gather( const orc_mesh* m, orc_grid* grid, int halo_width )
{
    return halo_exchange( m, grid, halo_width, 0 );
}

int orc_halo_scatter_sum( const orc_mesh* m, orc_grid* grid, int halo_width )
{
    return halo_exchange( m, grid, halo_width, 1 );
}

/* ------------------------------------------------------------------------ */
/* Global drivers                                                           */
/* ------------------------------------------------------------------------ */

static int check_p2g( const orc_mesh* m, int order, int op,
                      const orc_pfield* points, const orc_pfield* values,
                      const orc_grid* grid )
{
    if ( order < 0 || order > 6 || op < 0 || op > 4 )
        return 1;
    if ( points->ncomp != 3 || values->ncomp != p2g_point_comps( op ) ||
         values->n < points->n )
        return 1;
    return check_grid( m, grid, p2g_grid_dofs( op ) );
}

static int check_g2p( const orc_mesh* m, int order, int op,
                      const orc_pfield* points, const orc_pfield* values,
                      const orc_grid* grid )
{
    if ( order < 0 || order > 6 || op < 0 || op > 4 )
        return 1;
    if ( points->ncomp != 3 || values->ncomp != g2p_point_comps( op ) ||
         values->n < points->n )
        return 1;
    return check_grid( m, grid, g2p_grid_dofs( op ) );
}

static void load_point( const orc_pfield* points, int64_t p, double px[3] )
{
    for ( int d = 0; d < 3; ++d )
        px[d] = pf_get( points, p, d );
}

/* p2g, grid/src/Cabana_Grid_Interpolation.hpp:1217-1268 (Serial: plain
 * accumulation in point order, then halo scatter Sum). */
int orc_p2g( const orc_mesh* m, int order, int op, double multiplier,
             const orc_pfield* points, const orc_pfield* values,
             orc_grid* grid, int halo_width )
{
    if ( check_p2g( m, order, op, points, values, grid ) )
        return 1;
    for ( int64_t p = 0; p < points->n; ++p )
    {
        double px[3];
        load_point( points, p, px );
        orc_spline_data sd;
        orc_evaluate_spline( m, order, px, &sd );
        p2g_point( op, multiplier, values, p, &sd, grid );
    }
    return orc_halo_scatter_sum( m, grid, halo_width );
}

/* g2p, grid/src/Cabana_Grid_Interpolation.hpp:831-881 (halo gather first). */
int orc_g2p( const orc_mesh* m, int order, int op, double multiplier,
             orc_grid* grid, int halo_width, const orc_pfield* points,
             orc_pfield* values )
{
    if ( check_g2p( m, order, op, points, values, grid ) )
        return 1;
    int rc = orc_halo_gather( m, grid, halo_width );
    if ( rc )
        return rc;
    for ( int64_t p = 0; p < points->n; ++p )
    {
        double px[3];
        load_point( points, p, px );
        orc_spline_data sd;
        orc_evaluate_spline( m, order, px, &sd );
        g2p_point( op, multiplier, grid, p, &sd, values );
    }
    return 0;
}

int orc_p2g_omp( const orc_mesh* m, int order, int op, double multiplier,
                 const orc_pfield* points, const orc_pfield* values,
                 orc_grid* grid, int halo_width, int nthreads )
{
    if ( check_p2g( m, order, op, points, values, grid ) )
        return 1;
#ifdef _OPENMP
    if ( nthreads <= 0 )
        nthreads = omp_get_max_threads();
#else
    nthreads = 1;
#endif
    const int dofs = grid->ext[3];
    const int64_t nn = (int64_t)grid->ext[0] * grid->ext[1] * grid->ext[2];
    const int64_t ncopy = nn * dofs;
    double* copies = (double*)calloc( (size_t)ncopy * nthreads, sizeof( double ) );
    if ( !copies )
        return 4;
#pragma omp parallel num_threads( nthreads )
    {
        int t = 0;
#ifdef _OPENMP
        t = omp_get_thread_num();
#endif
        /* Private copy in (i,j,k,d) row-major order: a ScatterView duplicate. */
        orc_grid priv;
        priv.data = copies + (int64_t)t * ncopy;
        for ( int d = 0; d < 4; ++d )
            priv.ext[d] = grid->ext[d];
        priv.stride[3] = 1;
        priv.stride[2] = dofs;
        priv.stride[1] = (int64_t)dofs * grid->ext[2];
        priv.stride[0] = (int64_t)dofs * grid->ext[2] * grid->ext[1];
#pragma omp for schedule( static )
        for ( int64_t p = 0; p < points->n; ++p )
        {
            double px[3];
            load_point( points, p, px );
            orc_spline_data sd;
            orc_evaluate_spline( m, order, px, &sd );
            p2g_point( op, multiplier, values, p, &sd, &priv );
        }
        /* contribute(): sum the duplicates into the array. */
#pragma omp for schedule( static )
        for ( int64_t e = 0; e < nn; ++e )
        {
            int i = (int)( e / ( (int64_t)grid->ext[1] * grid->ext[2] ) );
            int j = (int)( ( e / grid->ext[2] ) % grid->ext[1] );
            int k = (int)( e % grid->ext[2] );
            for ( int l = 0; l < dofs; ++l )
            {
                double s = 0.0;
                for ( int c = 0; c < nthreads; ++c )
                    s += copies[(int64_t)c * ncopy + e * dofs + l];
                *gd_ptr( grid, i, j, k, l ) += s;
            }
        }
    }
    free( copies );
    return orc_halo_scatter_sum( m, grid, halo_width );
}

int orc_g2p_omp( const orc_mesh* m, int order, int op, double multiplier,
                 orc_grid* grid, int halo_width, const orc_pfield* points,
                 orc_pfield* values, int nthreads )
{
    if ( check_g2p( m, order, op, points, values, grid ) )
        return 1;
    int rc = orc_halo_gather( m, grid, halo_width );
    if ( rc )
        return rc;
#ifdef _OPENMP
    if ( nthreads <= 0 )
        nthreads = omp_get_max_threads();
#else
    nthreads = 1;
#endif
#pragma omp parallel for schedule( static ) num_threads( nthreads )
    for ( int64_t p = 0; p < points->n; ++p )
    {
        double px[3];
        load_point( points, p, px );
        orc_spline_data sd;
        orc_evaluate_spline( m, order, px, &sd );
        g2p_point( op, multiplier, grid, p, &sd, values );
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Adjoints                                                                 */
/* ------------------------------------------------------------------------ */

/* Weight W = wi wj wk of one stencil node, its position gradient dW[e] and
 * Hessian ddW[e][f] (all in the physical frame). */
static void node_derivs( const orc_spline_data* sd, const double h[3][ORC_MAX_KNOT],
                         int i, int j, int k, double* W, double dW[3],
                         double ddW[3][3] )
{
    const int n[3] = { i, j, k };
    double w[3], g[3], hh[3];
    for ( int d = 0; d < 3; ++d )
    {
        w[d] = sd->w[d][n[d]];
        g[d] = sd->g[d][n[d]];
        hh[d] = h[d][n[d]];
    }
    *W = w[0] * w[1] * w[2];
    dW[0] = g[0] * w[1] * w[2];
    dW[1] = w[0] * g[1] * w[2];
    dW[2] = w[0] * w[1] * g[2];
    for ( int e = 0; e < 3; ++e )
        for ( int f = 0; f < 3; ++f )
        {
            double v = 1.0;
            for ( int d = 0; d < 3; ++d )
            {
                int cnt = ( d == e ) + ( d == f );
                v *= cnt == 0 ? w[d] : ( cnt == 1 ? g[d] : hh[d] );
            }
            ddW[e][f] = v;
        }
}

static int copy_grid_dense( const orc_grid* src, orc_grid* dst )
{
    int64_t n = (int64_t)src->ext[0] * src->ext[1] * src->ext[2] * src->ext[3];
    dst->data = (double*)calloc( (size_t)n, sizeof( double ) );
    if ( !dst->data )
        return 4;
    for ( int d = 0; d < 4; ++d )
        dst->ext[d] = src->ext[d];
    dst->stride[3] = 1;
    dst->stride[2] = src->ext[3];
    dst->stride[1] = (int64_t)src->ext[3] * src->ext[2];
    dst->stride[0] = (int64_t)src->ext[3] * src->ext[2] * src->ext[1];
    return 0;
}

static int is_owned( const orc_mesh* m, int i, int j, int k )
{
    int idx[3] = { i, j, k };
    for ( int d = 0; d < 3; ++d )
    {
        int lo, hi;
        owned_bounds( m, d, &lo, &hi );
        if ( idx[d] < lo || idx[d] >= hi )
            return 0;
    }
    return 1;
}

int orc_p2g_adjoint( const orc_mesh* m, int order, int op, double multiplier,
                     const orc_pfield* points, const orc_pfield* values,
                     orc_grid* grid_bar, int halo_width,
                     orc_pfield* values_bar, orc_pfield* points_bar,
                     int nthreads )
{
    if ( check_p2g( m, order, op, points, values, grid_bar ) )
        return 1;
    /* Transpose of the fold: owned cotangents copied into their ghosts,
     * every other node zero. */
    orc_grid gh;
    if ( copy_grid_dense( grid_bar, &gh ) )
        return 4;
    const int dofs = grid_bar->ext[3];
    for ( int i = 0; i < gh.ext[0]; ++i )
        for ( int j = 0; j < gh.ext[1]; ++j )
            for ( int k = 0; k < gh.ext[2]; ++k )
                if ( is_owned( m, i, j, k ) )
                    for ( int l = 0; l < dofs; ++l )
                        *gd_ptr( &gh, i, j, k, l ) =
                            *gd_ptr( grid_bar, i, j, k, l );
    int rc = orc_halo_gather( m, &gh, halo_width );
    if ( rc )
    {
        free( gh.data );
        return rc;
    }
    const int ncomp = p2g_point_comps( op );
#ifdef _OPENMP
    if ( nthreads <= 0 )
        nthreads = 1;
#else
    nthreads = 1;
#endif
    /* Each particle writes only its own cotangents: a plain parallel loop. */
#pragma omp parallel for schedule( static ) num_threads( nthreads )
    for ( int64_t p = 0; p < points->n; ++p )
    {
        double px[3];
        load_point( points, p, px );
        orc_spline_data sd;
        orc_evaluate_spline( m, order, px, &sd );
        double hs[3][ORC_MAX_KNOT];
        for ( int d = 0; d < 3; ++d )
            orc_spline_hessian( order, sd.x[d], sd.rdx[d], hs[d] );
        double v[9];
        for ( int c = 0; c < ncomp; ++c )
            v[c] = pf_get( values, p, c );
        double vb[9] = { 0 }, xb[3] = { 0, 0, 0 };
        const int K = sd.num_knot;
        for ( int i = 0; i < K; ++i )
            for ( int j = 0; j < K; ++j )
                for ( int k = 0; k < K; ++k )
                {
                    double W, dW[3], ddW[3][3];
                    node_derivs( &sd, hs, i, j, k, &W, dW, ddW );
                    int a = sd.s[0][i], b = sd.s[1][j], c = sd.s[2][k];
                    double gb[3];
                    for ( int l = 0; l < dofs; ++l )
                        gb[l] = *gd_ptr( &gh, a, b, c, l );
                    switch ( op )
                    {
                    case ORC_P2G_SCALAR_VALUE:
                        vb[0] += W * gb[0];
                        for ( int e = 0; e < 3; ++e )
                            xb[e] += v[0] * dW[e] * gb[0];
                        break;
                    case ORC_P2G_VECTOR_VALUE:
                        for ( int d = 0; d < 3; ++d )
                        {
                            vb[d] += W * gb[d];
                            for ( int e = 0; e < 3; ++e )
                                xb[e] += v[d] * dW[e] * gb[d];
                        }
                        break;
                    case ORC_P2G_SCALAR_GRADIENT:
                        for ( int d = 0; d < 3; ++d )
                        {
                            vb[0] += dW[d] * gb[d];
                            for ( int e = 0; e < 3; ++e )
                                xb[e] += v[0] * ddW[e][d] * gb[d];
                        }
                        break;
                    case ORC_P2G_VECTOR_DIVERGENCE:
                        for ( int d = 0; d < 3; ++d )
                        {
                            vb[d] += dW[d] * gb[0];
                            for ( int e = 0; e < 3; ++e )
                                xb[e] += v[d] * ddW[e][d] * gb[0];
                        }
                        break;
                    default:
                        for ( int d0 = 0; d0 < 3; ++d0 )
                            for ( int d1 = 0; d1 < 3; ++d1 )
                            {
                                vb[3 * d0 + d1] += dW[d1] * gb[d0];
                                for ( int e = 0; e < 3; ++e )
                                    xb[e] += v[3 * d0 + d1] * ddW[e][d1] *
                                             gb[d0];
                            }
                        break;
                    }
                }
        if ( values_bar )
            for ( int c = 0; c < ncomp; ++c )
                *pf_ptr( values_bar, p, c ) += multiplier * vb[c];
        if ( points_bar )
            for ( int e = 0; e < 3; ++e )
                *pf_ptr( points_bar, p, e ) += multiplier * xb[e];
    }
    free( gh.data );
    return 0;
}

int orc_g2p_adjoint( const orc_mesh* m, int order, int op, double multiplier,
                     orc_grid* grid, int halo_width, const orc_pfield* points,
                     const orc_pfield* values_bar, orc_grid* grid_bar,
                     orc_pfield* points_bar, int nthreads )
{
    if ( check_g2p( m, order, op, points, values_bar, grid ) )
        return 1;
    if ( grid_bar && check_grid( m, grid_bar, g2p_grid_dofs( op ) ) )
        return 1;
    int rc = orc_halo_gather( m, grid, halo_width );
    if ( rc )
        return rc;
    const int dofs = grid->ext[3];
    orc_grid acc;
    if ( copy_grid_dense( grid, &acc ) )
        return 4;
    const int ncomp = g2p_point_comps( op );
#ifdef _OPENMP
    if ( nthreads <= 0 )
        nthreads = 1;
#else
    nthreads = 1;
#endif
    /* Grid cotangents from several threads go to private copies that are
     * summed afterwards (the ScatterView duplication of the forward p2g). */
    const int64_t ncopy = (int64_t)acc.ext[0] * acc.ext[1] * acc.ext[2] * dofs;
    double* copies = NULL;
    if ( nthreads > 1 )
    {
        copies = (double*)calloc( (size_t)ncopy * nthreads, sizeof( double ) );
        if ( !copies )
        {
            free( acc.data );
            return 4;
        }
    }
#pragma omp parallel num_threads( nthreads )
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        orc_grid mine = acc;
        if ( copies )
            mine.data = copies + (int64_t)tid * ncopy;
#pragma omp for schedule( static )
    for ( int64_t p = 0; p < points->n; ++p )
    {
        double px[3];
        load_point( points, p, px );
        orc_spline_data sd;
        orc_evaluate_spline( m, order, px, &sd );
        double hs[3][ORC_MAX_KNOT];
        for ( int d = 0; d < 3; ++d )
            orc_spline_hessian( order, sd.x[d], sd.rdx[d], hs[d] );
        double u[9];
        for ( int c = 0; c < ncomp; ++c )
            u[c] = pf_get( values_bar, p, c );
        double xb[3] = { 0, 0, 0 };
        const int K = sd.num_knot;
        for ( int i = 0; i < K; ++i )
            for ( int j = 0; j < K; ++j )
                for ( int k = 0; k < K; ++k )
                {
                    double W, dW[3], ddW[3][3];
                    node_derivs( &sd, hs, i, j, k, &W, dW, ddW );
                    int a = sd.s[0][i], b = sd.s[1][j], c = sd.s[2][k];
                    double g[3], gb[3] = { 0, 0, 0 };
                    for ( int l = 0; l < dofs; ++l )
                        g[l] = *gd_ptr( grid, a, b, c, l );
                    switch ( op )
                    {
                    case ORC_G2P_SCALAR_VALUE:
                        gb[0] = u[0] * W;
                        for ( int e = 0; e < 3; ++e )
                            xb[e] += u[0] * dW[e] * g[0];
                        break;
                    case ORC_G2P_VECTOR_VALUE:
                        for ( int d = 0; d < 3; ++d )
                        {
                            gb[d] = u[d] * W;
                            for ( int e = 0; e < 3; ++e )
                                xb[e] += u[d] * dW[e] * g[d];
                        }
                        break;
                    case ORC_G2P_SCALAR_GRADIENT:
                        for ( int d = 0; d < 3; ++d )
                        {
                            gb[0] += u[d] * dW[d];
                            for ( int e = 0; e < 3; ++e )
                                xb[e] += u[d] * ddW[e][d] * g[0];
                        }
                        break;
                    case ORC_G2P_VECTOR_GRADIENT:
                        for ( int d0 = 0; d0 < 3; ++d0 )
                            for ( int d1 = 0; d1 < 3; ++d1 )
                            {
                                gb[d1] += u[3 * d0 + d1] * dW[d0];
                                for ( int e = 0; e < 3; ++e )
                                    xb[e] += u[3 * d0 + d1] * ddW[e][d0] *
                                             g[d1];
                            }
                        break;
                    default:
                        for ( int d = 0; d < 3; ++d )
                        {
                            gb[d] = u[0] * dW[d];
                            for ( int e = 0; e < 3; ++e )
                                xb[e] += u[0] * ddW[e][d] * g[d];
                        }
                        break;
                    }
                    for ( int l = 0; l < dofs; ++l )
                        *gd_ptr( &mine, a, b, c, l ) += multiplier * gb[l];
                }
        if ( points_bar )
            for ( int e = 0; e < 3; ++e )
                *pf_ptr( points_bar, p, e ) += multiplier * xb[e];
    }
        if ( copies )
        {
#pragma omp for schedule( static )
            for ( int64_t e = 0; e < ncopy; ++e )
            {
                double s = 0.0;
                for ( int c = 0; c < nthreads; ++c )
                    s += copies[(int64_t)c * ncopy + e];
                acc.data[e] += s;
            }
        }
    }
    free( copies );
    /* Transpose of the ghost fill: fold ghosts into owners. */
    rc = orc_halo_scatter_sum( m, &acc, halo_width );
    if ( !rc && grid_bar )
        for ( int i = 0; i < acc.ext[0]; ++i )
            for ( int j = 0; j < acc.ext[1]; ++j )
                for ( int k = 0; k < acc.ext[2]; ++k )
                    if ( is_owned( m, i, j, k ) )
                        for ( int l = 0; l < dofs; ++l )
                            *gd_ptr( grid_bar, i, j, k, l ) +=
                                *gd_ptr( &acc, i, j, k, l );
    free( acc.data );
    return rc;
}

/* ------------------------------------------------------------------------ */
/* Binning                                                                  */
/* ------------------------------------------------------------------------ */

/* cellsBetween, CartesianGrid.hpp:279-283 */
static int cells_between( double max, double min, double rdelta )
{
    return (int)floor( ( max - min ) * rdelta );
}

/* CartesianGrid::init, CartesianGrid.hpp:61-73 */
void orc_cart_grid_init( orc_cart_grid* g, const double min[3],
                         const double max[3], const double delta[3] )
{
    for ( int d = 0; d < 3; ++d )
    {
        g->min[d] = min[d];
        g->max[d] = max[d];
        g->nx[d] = cells_between( max[d], min[d], 1.0 / delta[d] );
        g->dx[d] = ( max[d] - min[d] ) / g->nx[d];
        g->rdx[d] = 1.0 / g->dx[d];
    }
}

/* locatePoint, CartesianGrid.hpp:171-182 (clamp only at == nx) */
void orc_cart_locate( const orc_cart_grid* g, const double p[3], int ijk[3] )
{
    for ( int d = 0; d < 3; ++d )
    {
        int c = cells_between( p[d], g->min[d], g->rdx[d] );
        ijk[d] = ( c == g->nx[d] ) ? c - 1 : c;
    }
}

/* cardinalCellIndex, CartesianGrid.hpp:226-231 (i slowest, k fastest) */
int64_t orc_cart_cardinal( const orc_cart_grid* g, const int ijk[3] )
{
    return ( (int64_t)ijk[0] * g->nx[1] + ijk[1] ) * g->nx[2] + ijk[2];
}

int orc_lcl_build( const orc_pfield* positions, int64_t begin, int64_t end,
                   const double delta[3], const double min[3],
                   const double max[3], int32_t* counts, int64_t* offsets,
                   int64_t* permutes )
{
    if ( positions->ncomp != 3 || begin < 0 || end < begin ||
         end > positions->n )
        return 1;
    orc_cart_grid g;
    orc_cart_grid_init( &g, min, max, delta );
    int64_t ncell = (int64_t)g.nx[0] * g.nx[1] * g.nx[2];
    int64_t* cell = (int64_t*)malloc( sizeof( int64_t ) * ( end - begin + 1 ) );
    if ( !cell )
        return 4;
    for ( int64_t c = 0; c < ncell; ++c )
        counts[c] = 0;
    /* Count, LinkedCellList.hpp:682-699. */
    for ( int64_t p = begin; p < end; ++p )
    {
        double px[3];
        load_point( positions, p, px );
        int ijk[3];
        orc_cart_locate( &g, px, ijk );
        for ( int d = 0; d < 3; ++d )
            if ( ijk[d] < 0 || ijk[d] >= g.nx[d] )
            {
                free( cell );
                return 5;
            }
        cell[p - begin] = orc_cart_cardinal( &g, ijk );
        counts[cell[p - begin]] += 1;
    }
    /* Exclusive scan, :701-712. */
    int64_t run = 0;
    for ( int64_t c = 0; c < ncell; ++c )
    {
        offsets[c] = run;
        run += counts[c];
    }
    /* Reset counts and place in ascending particle order (Serial), :714-731. */
    for ( int64_t c = 0; c < ncell; ++c )
        counts[c] = 0;
    for ( int64_t p = begin; p < end; ++p )
    {
        int64_t cid = cell[p - begin];
        int32_t slot = counts[cid]++;
        permutes[offsets[cid] + slot] = p;
    }
    free( cell );
    return 0;
}

int orc_permute( const int64_t* permutes, int64_t begin, int64_t end,
                 orc_pfield* field )
{
    if ( begin < 0 || end < begin || end > field->n )
        return 1;
    int64_t n = end - begin;
    int nc = field->ncomp;
    double* scratch = (double*)malloc( sizeof( double ) * ( n * nc + 1 ) );
    if ( !scratch )
        return 4;
    for ( int64_t i = begin; i < end; ++i )
    {
        int64_t src = permutes[i - begin];
        for ( int c = 0; c < nc; ++c )
            scratch[( i - begin ) * nc + c] = pf_get( field, src, c );
    }
    for ( int64_t i = begin; i < end; ++i )
        for ( int c = 0; c < nc; ++c )
            *pf_ptr( field, i, c ) = scratch[( i - begin ) * nc + c];
    free( scratch );
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Particle migration on one rank                                           */
/* ------------------------------------------------------------------------ */

/* particleMigrate (grid/src/Cabana_Grid_ParticleDistributor.hpp:330-352) on
 * one rank, over row-major positions x[p][3].  migrateCount (:170-221):
 * with force == 0 and no particle outside the ghosted corners moved in by
 * min_halo_width cells, returns -1 and changes nothing.  Otherwise
 * getMigrateDestinations (:87-148): the neighbour index of each dimension
 * is 0 below the owned low corner, 2 above the owned high corner, 1 inside;
 * getTopology (:48-57) on one rank gives this rank in every direction of a
 * periodic dimension and -1 past a non-periodic boundary, so keep[p] = 0
 * exactly when some non-periodic dimension is out; then periodic
 * coordinates above the global high corner lose the extent, those below the
 * low corner gain it.  Returns the number of particles kept. */
int64_t orc_particle_migrate( const orc_mesh* m, double* x, int64_t n, int min_halo_width,
                              int force, int32_t* keep )
{
    double own_lo[3], own_hi[3], band_lo[3], band_hi[3], extent[3];
    for ( int d = 0; d < 3; ++d )
    {
        /* LocalMesh::lowCorner/highCorner, Cabana_Grid_LocalMesh.hpp:88-107 */
        own_lo[d] = m->global_low[d] + m->cell_size[d] * 0;
        own_hi[d] = m->global_low[d] + m->cell_size[d] * ( 0 + m->num_cell[d] );
        double ghost_lo = own_lo[d] - m->halo * m->cell_size[d];
        double ghost_hi = own_hi[d] + m->halo * m->cell_size[d];
        band_lo[d] = ghost_lo + min_halo_width * m->cell_size[d];
        band_hi[d] = ghost_hi - min_halo_width * m->cell_size[d];
        /* GlobalMesh::extent, Cabana_Grid_GlobalMesh.hpp:153-157 */
        extent[d] = m->global_high[d] - m->global_low[d];
    }
    if ( !force )
    {
        int64_t count = 0;
        for ( int64_t p = 0; p < n; ++p )
            for ( int d = 0; d < 3; ++d )
                if ( x[3 * p + d] < band_lo[d] || x[3 * p + d] > band_hi[d] )
                {
                    count += 1;
                    break;
                }
        if ( count == 0 )
            return -1;
    }
    int64_t kept = 0;
    for ( int64_t p = 0; p < n; ++p )
    {
        int stays = 1;
        for ( int d = 0; d < 3; ++d )
        {
            double* xp = &x[3 * p + d];
            int nid = 1;
            if ( *xp < own_lo[d] )
                nid = 0;
            else if ( *xp > own_hi[d] )
                nid = 2;
            if ( nid != 1 && !m->periodic[d] )
                stays = 0;
            if ( m->periodic[d] )
            {
                if ( *xp > m->global_high[d] )
                    *xp -= extent[d];
                else if ( *xp < m->global_low[d] )
                    *xp += extent[d];
            }
        }
        keep[p] = stays;
        kept += stays;
    }
    return kept;
}

/* ------------------------------------------------------------------------ */
/* BOV output on one rank                                                    */
/* ------------------------------------------------------------------------ */

/* Experimental::BovWriter::writeTimeStep (grid/src/Cabana_Grid_BovWriter.hpp
 * :180-338) on one rank.  entity 0 = Node (grid extents num_node), 1 = Cell
 * (num_cell + 2h).  Global extents are the cell counts, +1 for nodes
 * (:196-203); the owned extents of the single (last) block are the same
 * (:210-225); a node array with a periodic dimension is first gathered with
 * a width-0 halo (:227-240).  The data file holds grid(h + i, h + j, h + k, l)
 * in (k, j, i, l) order (reorderView :140-152, C-order subarray :82-122);
 * the header lines are those of :276-337, doubles in the stream's default
 * format, which is printf's %g. */
int orc_bov_write( const orc_mesh* m, orc_grid* grid, int entity, const char* label,
                   const char* prefix, int time_step_index, double time, int gather_array )
{
    const int node = entity == 0;
    long ext[3];
    for ( int d = 0; d < 3; ++d )
    {
        ext[d] = m->num_cell[d] + ( node ? 1 : 0 );
        int want = node ? m->num_node[d] : m->num_cell[d] + 2 * m->halo;
        if ( grid->ext[d] != want )
            return 1;
    }
    if ( node && gather_array )
        for ( int d = 0; d < 3; ++d )
            if ( m->periodic[d] )
            {
                if ( orc_halo_gather( m, grid, 0 ) )
                    return 1;
                break;
            }
    char base[4096];
    snprintf( base, sizeof( base ), "%s_%06d", prefix, time_step_index );
    char data_name[4200], header_name[4200];
    snprintf( data_name, sizeof( data_name ), "%s.dat", base );
    snprintf( header_name, sizeof( header_name ), "%s.bov", base );
    FILE* f = fopen( data_name, "wb" );
    if ( !f )
        return 2;
    const int h = m->halo;
    for ( long k = 0; k < ext[2]; ++k )
        for ( long j = 0; j < ext[1]; ++j )
            for ( long i = 0; i < ext[0]; ++i )
                for ( int l = 0; l < grid->ext[3]; ++l )
                {
                    double v = *gd_ptr( grid, h + (int)i, h + (int)j, h + (int)k, l );
                    fwrite( &v, sizeof( double ), 1, f );
                }
    fclose( f );
    f = fopen( header_name, "w" );
    if ( !f )
        return 2;
    fprintf( f, "TIME: %g\n", time );
    fprintf( f, "DATA_FILE: %s\n", data_name );
    fprintf( f, "DATA_SIZE: %ld %ld %ld \n", ext[0], ext[1], ext[2] );
    fprintf( f, "DATA_FORMAT: DOUBLE\n" );
    fprintf( f, "VARIABLE: %s\n", label );
    fprintf( f, "DATA_ENDIAN: LITTLE\n" );
    fprintf( f, "CENTERING: %s\n", node ? "nodal" : "zonal" );
    fprintf( f, "BRICK_ORIGIN: %g %g %g \n", m->global_low[0], m->global_low[1],
             m->global_low[2] );
    fprintf( f, "BRICK_SIZE: %g %g %g \n", m->num_cell[0] * m->cell_size[0],
             m->num_cell[1] * m->cell_size[1], m->num_cell[2] * m->cell_size[2] );
    fprintf( f, "DATA_COMPONENTS: %d\n", grid->ext[3] );
    fclose( f );
    return 0;
}
