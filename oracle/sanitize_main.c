/*
 * Sanitizer driver for the oracle (test infrastructure only): built with
 * AddressSanitizer and UndefinedBehaviorSanitizer by `make -C oracle
 * sanitize`, it runs every oracle function on small seeded cases (splines of
 * orders 0 to 6, P2G / G2P of every operator on every mesh entity, Serial
 * and OpenMP, both adjoints, halos of every width, binning, permute,
 * particle migration and BOV output) and checks a few invariants, so that an
 * out-of-bounds access or undefined operation anywhere in the oracle stops
 * it with a report.  Exit status 0 when everything ran clean.
 */
#include "oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static uint64_t rng_state = 0x9e3779b97f4a7c15ull;

/* splitmix64 */
static double uniform01( void )
{
    uint64_t z = ( rng_state += 0x9e3779b97f4a7c15ull );
    z = ( z ^ ( z >> 30 ) ) * 0xbf58476d1ce4e5b9ull;
    z = ( z ^ ( z >> 27 ) ) * 0x94d049bb133111ebull;
    z ^= z >> 31;
    return (double)( z >> 11 ) * ( 1.0 / 9007199254740992.0 );
}

static int failures = 0;

static void check( int ok, const char* what )
{
    if ( !ok )
    {
        fprintf( stderr, "check failed: %s\n", what );
        ++failures;
    }
}

static orc_pfield pf( double* data, int64_t n, int ncomp )
{
    orc_pfield f;
    f.data = data;
    f.n = n;
    f.ncomp = ncomp;
    f.vl = 1;
    f.soa_stride = ncomp;
    return f;
}

static double* grid_alloc( const orc_mesh* m, int dofs, orc_grid* g )
{
    const int64_t nn = (int64_t)m->num_node[0] * m->num_node[1] * m->num_node[2] * dofs;
    double* d = (double*)calloc( (size_t)nn, sizeof( double ) );
    g->data = d;
    g->ext[0] = m->num_node[0];
    g->ext[1] = m->num_node[1];
    g->ext[2] = m->num_node[2];
    g->ext[3] = dofs;
    g->stride[3] = 1;
    g->stride[2] = dofs;
    g->stride[1] = (int64_t)dofs * m->num_node[2];
    g->stride[0] = (int64_t)dofs * m->num_node[2] * m->num_node[1];
    return d;
}

static const int p2g_comps[5][2] = { { 1, 1 }, { 3, 3 }, { 1, 3 }, { 3, 1 }, { 9, 3 } };
static const int g2p_comps[5][2] = { { 1, 1 }, { 3, 3 }, { 3, 1 }, { 9, 3 }, { 1, 3 } };

static void splines( void )
{
    for ( int order = 0; order <= 6; ++order )
        for ( int s = 0; s < 50; ++s )
        {
            double xp = -3.0 + 9.0 * uniform01(), w[ORC_MAX_KNOT], g[ORC_MAX_KNOT],
                   h[ORC_MAX_KNOT];
            int idx[ORC_MAX_KNOT];
            const double x0 = orc_spline_map( order, xp, 3.7, -5.0 );
            orc_spline_stencil( order, x0, idx );
            orc_spline_value( order, x0, w );
            orc_spline_gradient( order, x0, 3.7, g );
            orc_spline_hessian( order, x0, 3.7, h );
            double sw = 0.0, sg = 0.0;
            for ( int n = 0; n < orc_spline_num_knot( order ); ++n )
            {
                sw += w[n];
                sg += g[n];
            }
            check( fabs( sw - 1.0 ) < 1e-12 && fabs( sg ) < 1e-9, "spline partition of unity" );
        }
}

static void interpolation( int entity, int order, const int periodic[3] )
{
    const double low[3] = { 0.3, -1.1, 2.0 };
    const double cell[3] = { 0.37, 0.37, 0.37 };
    const double high[3] = { low[0] + 5 * cell[0], low[1] + 4 * cell[1], low[2] + 6 * cell[2] };
    const int halo = ( order / 2 > 1 ? order / 2 : 1 ) + 1;
    orc_mesh m;
    check( orc_mesh_init( &m, low, high, cell, halo, periodic ) == 0, "mesh init" );
    check( orc_mesh_set_entity( &m, entity ) == 0, "mesh entity" );
    const int n = 40;
    double x[40 * 3];
    for ( int p = 0; p < n; ++p )
        for ( int d = 0; d < 3; ++d )
            x[3 * p + d] = low[d] + uniform01() * ( high[d] - low[d] );
    orc_pfield pts = pf( x, n, 3 );
    for ( int op = 0; op < 5; ++op )
    {
        double vals[40 * 9], out[40 * 9], vbar[40 * 9], xbar[40 * 3];
        for ( int i = 0; i < 40 * 9; ++i )
            vals[i] = uniform01() - 0.5;
        orc_grid g, gb;
        double* gd = grid_alloc( &m, p2g_comps[op][1], &g );
        orc_pfield v = pf( vals, n, p2g_comps[op][0] );
        check( orc_p2g( &m, order, op, -0.5, &pts, &v, &g, -1 ) == 0, "p2g" );
        check( orc_p2g_omp( &m, order, op, -0.5, &pts, &v, &g, -1, 3 ) == 0, "p2g omp" );
        double* gbd = grid_alloc( &m, p2g_comps[op][1], &gb );
        for ( int64_t i = 0; i < (int64_t)gb.ext[0] * gb.ext[1] * gb.ext[2] * gb.ext[3]; ++i )
            gbd[i] = uniform01();
        memset( vbar, 0, sizeof( vbar ) );
        memset( xbar, 0, sizeof( xbar ) );
        orc_pfield vb = pf( vbar, n, p2g_comps[op][0] ), xb = pf( xbar, n, 3 );
        check( orc_p2g_adjoint( &m, order, op, 0.7, &pts, &v, &gb, -1, &vb, &xb, 2 ) == 0,
               "p2g adjoint" );
        free( gd );
        free( gbd );

        gd = grid_alloc( &m, g2p_comps[op][1], &g );
        for ( int64_t i = 0; i < (int64_t)g.ext[0] * g.ext[1] * g.ext[2] * g.ext[3]; ++i )
            gd[i] = uniform01();
        memset( out, 0, sizeof( out ) );
        orc_pfield o = pf( out, n, g2p_comps[op][0] );
        check( orc_g2p( &m, order, op, 1.3, &g, -1, &pts, &o ) == 0, "g2p" );
        check( orc_g2p_omp( &m, order, op, 1.3, &g, -1, &pts, &o, 3 ) == 0, "g2p omp" );
        gbd = grid_alloc( &m, g2p_comps[op][1], &gb );
        memset( xbar, 0, sizeof( xbar ) );
        orc_pfield ub = pf( out, n, g2p_comps[op][0] );
        check( orc_g2p_adjoint( &m, order, op, 1.1, &g, -1, &pts, &ub, &gb, &xb, 2 ) == 0,
               "g2p adjoint" );
        for ( int w = 0; w <= halo; ++w )
        {
            check( orc_halo_gather( &m, &g, w ) == 0, "halo gather" );
            check( orc_halo_scatter_sum( &m, &g, w ) == 0, "halo scatter" );
        }
        free( gd );
        free( gbd );
    }
}

static void binning( void )
{
    const int n = 3000;
    double* x = (double*)malloc( sizeof( double ) * 3 * n );
    double* y = (double*)malloc( sizeof( double ) * 3 * n );
    for ( int i = 0; i < 3 * n; ++i )
        x[i] = 4.0 * uniform01();
    x[0] = 4.0; /* on the upper boundary: clamped into the last bin */
    memcpy( y, x, sizeof( double ) * 3 * n );
    const double delta[3] = { 0.5, 0.5, 0.25 }, mn[3] = { 0, 0, 0 }, mx[3] = { 4, 4, 4 };
    const int nb = 8 * 8 * 16;
    int32_t* counts = (int32_t*)malloc( sizeof( int32_t ) * nb );
    int64_t* offsets = (int64_t*)malloc( sizeof( int64_t ) * nb );
    int64_t* perm = (int64_t*)malloc( sizeof( int64_t ) * n );
    orc_pfield pos = pf( x, n, 3 );
    check( orc_lcl_build( &pos, 100, n, delta, mn, mx, counts, offsets, perm ) == 0, "lcl build" );
    int64_t total = 0;
    for ( int c = 0; c < nb; ++c )
        total += counts[c];
    check( total == n - 100, "lcl counts" );
    check( orc_permute( perm, 100, n, &pos ) == 0, "permute" );
    /* migration on a mixed-periodic mesh */
    orc_mesh m;
    const double low[3] = { 0, 0, 0 }, high[3] = { 4, 4, 4 }, cs[3] = { 0.5, 0.5, 0.5 };
    const int per[3] = { 1, 0, 1 };
    check( orc_mesh_init( &m, low, high, cs, 1, per ) == 0, "mesh" );
    for ( int i = 0; i < 3 * n; ++i )
        y[i] = -3.0 + 10.0 * uniform01();
    int32_t* keep = (int32_t*)malloc( sizeof( int32_t ) * n );
    const int64_t kept = orc_particle_migrate( &m, y, n, 1, 1, keep );
    check( kept >= 0 && kept <= n, "migrate" );
    free( keep );
    free( perm );
    free( offsets );
    free( counts );
    free( y );
    free( x );
}

static void bov( void )
{
    orc_mesh m;
    const double low[3] = { 1.2, 3.3, -2.8 }, cs[3] = { 0.23, 0.23, 0.23 };
    const double high[3] = { low[0] + 0.23 * 5, low[1] + 0.23 * 4, low[2] + 0.23 * 3 };
    const int per[3] = { 1, 1, 0 };
    check( orc_mesh_init( &m, low, high, cs, 2, per ) == 0, "mesh" );
    orc_grid g;
    double* gd = grid_alloc( &m, 3, &g );
    check( orc_bov_write( &m, &g, 0, "f", "/tmp/orc_sanitize_bov", 3, 0.5, 1 ) == 0, "bov node" );
    free( gd );
    orc_mesh mc = m;
    check( orc_mesh_set_entity( &mc, 1 ) == 0, "cell entity" );
    gd = grid_alloc( &mc, 2, &g );
    check( orc_bov_write( &m, &g, 1, "c", "/tmp/orc_sanitize_bov_cell", 4, 1.5, 1 ) == 0,
           "bov cell" );
    free( gd );
    remove( "/tmp/orc_sanitize_bov_000003.dat" );
    remove( "/tmp/orc_sanitize_bov_000003.bov" );
    remove( "/tmp/orc_sanitize_bov_cell_000004.dat" );
    remove( "/tmp/orc_sanitize_bov_cell_000004.bov" );
}

int main( void )
{
    splines();
    const int periodic[2][3] = { { 1, 1, 1 }, { 1, 0, 1 } };
    for ( int e = 0; e < 8; ++e )
        for ( int order = 0; order <= 6; ++order )
            for ( int q = 0; q < 2; ++q )
                if ( e == 0 || ( order >= 1 && order <= 3 ) )
                    interpolation( e, order, periodic[q] );
    binning();
    bov();
    if ( failures )
    {
        fprintf( stderr, "%d checks failed\n", failures );
        return 1;
    }
    printf( "oracle sanitizer run: clean\n" );
    return 0;
}
