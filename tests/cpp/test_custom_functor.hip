// User-written functors through include/cabana_amd_custom.hpp on the GPU.
// Every built-in operator is written again as a custom functor with the
// P2G:: / G2P:: helpers, the way a Cabana user writes one, and run through
// the custom point loop; its grids and particle results must match the
// library's own kernels (cg::createScalarValueP2G and the rest) to rounding,
// for Spline<1> to Spline<3>, and a functor combining two helpers must match
// the two built-in calls.  Exit status 0 and "ALL PASSED" mean success.
#include "cabana_amd_custom.hpp"

#include <random>

namespace cg = cabana_amd::grid;

#include "standins.hpp"

// --- custom P2G functors (Cabana_Grid_Interpolation.hpp:1324-1566 written
// by hand) --------------------------------------------------------------------

struct ScalarValueP2G
{
    const double* x;
    double m;
    template <class SD, class V>
    __device__ void operator()( const SD& sd, int p, const V& v ) const
    {
        cg::P2G::value( m * x[p], sd, v );
    }
};
struct VectorValueP2G
{
    const double* x;
    double m;
    template <class SD, class V>
    __device__ void operator()( const SD& sd, int p, const V& v ) const
    {
        const double d[3] = { m * x[3 * p], m * x[3 * p + 1], m * x[3 * p + 2] };
        cg::P2G::value( d, sd, v );
    }
};
struct ScalarGradientP2G
{
    const double* x;
    double m;
    template <class SD, class V>
    __device__ void operator()( const SD& sd, int p, const V& v ) const
    {
        cg::P2G::gradient( m * x[p], sd, v );
    }
};
struct VectorDivergenceP2G
{
    const double* x;
    double m;
    template <class SD, class V>
    __device__ void operator()( const SD& sd, int p, const V& v ) const
    {
        const double d[3] = { m * x[3 * p], m * x[3 * p + 1], m * x[3 * p + 2] };
        cg::P2G::divergence( d, sd, v );
    }
};
struct TensorDivergenceP2G
{
    const double* x;
    double m;
    template <class SD, class V>
    __device__ void operator()( const SD& sd, int p, const V& v ) const
    {
        double t[3][3];
        for ( int a = 0; a < 3; ++a )
            for ( int b = 0; b < 3; ++b )
                t[a][b] = m * x[9 * p + 3 * a + b];
        cg::P2G::divergence( t, sd, v );
    }
};
// Two helper calls in one functor: the gradients of x and of 2x add up to
// the built-in ScalarGradientP2G with multiplier 3.
struct TwoGradients
{
    const double* x;
    template <class SD, class V>
    __device__ void operator()( const SD& sd, int p, const V& v ) const
    {
        cg::P2G::gradient( x[p], sd, v );
        cg::P2G::gradient( 2.0 * x[p], sd, v );
    }
};

// --- custom G2P functors ------------------------------------------------------

struct ScalarValueG2P
{
    double* out;
    double m;
    template <class SD, class V>
    __device__ void operator()( const SD& sd, int p, const V& v ) const
    {
        double r;
        cg::G2P::value( v, sd, r );
        out[p] += m * r;
    }
};
struct VectorValueG2P
{
    double* out;
    double m;
    template <class SD, class V>
    __device__ void operator()( const SD& sd, int p, const V& v ) const
    {
        double r[3];
        cg::G2P::value( v, sd, r );
        for ( int d = 0; d < 3; ++d )
            out[3 * p + d] += m * r[d];
    }
};
struct ScalarGradientG2P
{
    double* out;
    double m;
    template <class SD, class V>
    __device__ void operator()( const SD& sd, int p, const V& v ) const
    {
        double r[3];
        cg::G2P::gradient( v, sd, r );
        for ( int d = 0; d < 3; ++d )
            out[3 * p + d] += m * r[d];
    }
};
struct VectorGradientG2P
{
    double* out;
    double m;
    template <class SD, class V>
    __device__ void operator()( const SD& sd, int p, const V& v ) const
    {
        double r[3][3];
        cg::G2P::gradient( v, sd, r );
        for ( int a = 0; a < 3; ++a )
            for ( int b = 0; b < 3; ++b )
                out[9 * p + 3 * a + b] += m * r[a][b];
    }
};
struct VectorDivergenceG2P
{
    double* out;
    double m;
    template <class SD, class V>
    __device__ void operator()( const SD& sd, int p, const V& v ) const
    {
        double r;
        cg::G2P::divergence( v, sd, r );
        out[p] += m * r;
    }
};

// --- checks -------------------------------------------------------------------

static void expect_close( const std::vector<double>& a, const std::vector<double>& b,
                          const char* what, double rtol )
{
    double scale = 0.0, err = 0.0;
    for ( std::size_t i = 0; i < a.size(); ++i )
    {
        scale = std::fmax( scale, std::fabs( b[i] ) );
        err = std::fmax( err, std::fabs( a[i] - b[i] ) );
    }
    ++g_checks;
    if ( !( err <= rtol * std::fmax( scale, 1e-300 ) ) || a.size() != b.size() )
    {
        ++g_failures;
        std::printf( "FAIL %s: max |diff| %.3e against max |ref| %.3e\n", what, err, scale );
    }
}

struct Setup
{
    GlobalMesh gm{ { 0.3, -1.1, 2.0 }, { 0.3 + 0.25 * 10, -1.1 + 0.25 * 9, 2.0 + 0.25 * 11 },
                   0.25 };
    int n[3] = { 10, 9, 11 };
    int h;
    std::shared_ptr<LocalGrid> lg;
    int num_point = 4000;
    DeviceView<double, 2> points;
    DeviceView<double, 3> tensor;
    std::vector<double> th;

    explicit Setup( int halo ) : h( halo )
    {
        lg = std::make_shared<LocalGrid>( LocalGrid{ GlobalGrid{ gm }, h } );
        std::mt19937_64 rng( 1234 + halo );
        std::uniform_real_distribution<double> u( 0.0, 1.0 );
        std::normal_distribution<double> g( 0.0, 1.0 );
        std::vector<double> ph( 3 * num_point );
        for ( int p = 0; p < num_point; ++p )
            for ( int d = 0; d < 3; ++d )
                ph[3 * p + d] = gm.lo[d] + u( rng ) * ( gm.hi[d] - gm.lo[d] );
        points = DeviceView<double, 2>( { std::size_t( num_point ), 3 } );
        points.from_host( ph );
        th.resize( 9 * num_point );
        for ( auto& x : th )
            x = g( rng );
        tensor = DeviceView<double, 3>( { std::size_t( num_point ), 3, 3 } );
        tensor.from_host( th );
    }
    template <int R>
    DeviceView<double, R> field( std::array<std::size_t, R> ext ) const
    {
        DeviceView<double, R> v( ext );
        std::vector<double> h( v.size() );
        for ( std::size_t i = 0; i < h.size(); ++i )
            h[i] = th[i % th.size()] * 0.7 + 0.1 * double( i % 7 );
        v.from_host( h );
        return v;
    }
    std::shared_ptr<Array<double>> grid( int dofs, bool random ) const
    {
        int nn[3] = { n[0], n[1], n[2] };
        auto a = make_array<double>( lg, dofs, nn );
        if ( random )
        {
            std::vector<double> h( a->v.size() );
            std::mt19937_64 rng( 99 + dofs );
            std::normal_distribution<double> g( 0.0, 1.0 );
            for ( auto& x : h )
                x = g( rng );
            a->v.from_host( h );
        }
        return a;
    }
};

template <int Order>
static void run_order()
{
    Setup s( Order / 2 + 1 );
    const std::size_t np = std::size_t( s.num_point );
    auto sc = s.field<1>( { np } );
    auto vec = s.field<2>( { np, 3 } );
    const double m = -0.6;
    const double tol = 1e-12;
    char what[96];

    // P2G: each operator as a custom functor against the built-in one.
    auto p2g_pair = [&]( int dofs, const char* name, auto custom, auto builtin ) {
        auto a = s.grid( dofs, false );
        auto b = s.grid( dofs, false );
        auto halo = cg::createHalo( cg::NodeHaloPattern<3>(), s.h, *a );
        cg::p2g( custom, s.points, np, cg::Spline<Order>(), *halo, *a );
        cg::p2g( builtin, s.points, np, cg::Spline<Order>(), *halo, *b );
        std::snprintf( what, sizeof( what ), "p2g %s order %d", name, Order );
        expect_close( a->v.to_host(), b->v.to_host(), what, tol );
    };
    p2g_pair( 1, "ScalarValue", ScalarValueP2G{ sc.data(), m }, cg::createScalarValueP2G( sc, m ) );
    p2g_pair( 3, "VectorValue", VectorValueP2G{ vec.data(), m },
              cg::createVectorValueP2G( vec, m ) );
    p2g_pair( 3, "ScalarGradient", ScalarGradientP2G{ sc.data(), m },
              cg::createScalarGradientP2G( sc, m ) );
    p2g_pair( 1, "VectorDivergence", VectorDivergenceP2G{ vec.data(), m },
              cg::createVectorDivergenceP2G( vec, m ) );
    p2g_pair( 3, "TensorDivergence", TensorDivergenceP2G{ s.tensor.data(), m },
              cg::createTensorDivergenceP2G( s.tensor, m ) );
    p2g_pair( 3, "two helpers in one functor", TwoGradients{ sc.data() },
              cg::createScalarGradientP2G( sc, 3.0 ) );

    // G2P: each operator reads a random grid (ghosts filled by the call).
    auto g2p_pair = [&]( int dofs, std::size_t ncomp, const char* name, auto make_custom,
                         auto make_builtin ) {
        auto grid = s.grid( dofs, true );
        auto halo = cg::createHalo( cg::NodeHaloPattern<3>(), s.h, *grid );
        DeviceView<double, 2> a( { np, ncomp } ), b( { np, ncomp } );
        cg::g2p( *grid, *halo, s.points, np, cg::Spline<Order>(), make_custom( a.data() ) );
        cg::g2p( *grid, *halo, s.points, np, cg::Spline<Order>(), make_builtin( b ) );
        std::snprintf( what, sizeof( what ), "g2p %s order %d", name, Order );
        expect_close( a.to_host(), b.to_host(), what, tol );
    };
    g2p_pair( 1, 1, "ScalarValue", [&]( double* o ) { return ScalarValueG2P{ o, m }; },
              [&]( DeviceView<double, 2>& o ) {
                  DeviceView<double, 1> v;
                  v.mem = o.mem;
                  v.ext[0] = o.ext[0];
                  v.str[0] = 1;
                  return cg::createScalarValueG2P( v, m );
              } );
    g2p_pair( 3, 3, "VectorValue", [&]( double* o ) { return VectorValueG2P{ o, m }; },
              [&]( DeviceView<double, 2>& o ) { return cg::createVectorValueG2P( o, m ); } );
    g2p_pair( 1, 3, "ScalarGradient", [&]( double* o ) { return ScalarGradientG2P{ o, m }; },
              [&]( DeviceView<double, 2>& o ) { return cg::createScalarGradientG2P( o, m ); } );
    g2p_pair( 3, 9, "VectorGradient", [&]( double* o ) { return VectorGradientG2P{ o, m }; },
              [&]( DeviceView<double, 2>& o ) {
                  DeviceView<double, 3> v;
                  v.mem = o.mem;
                  v.ext[0] = o.ext[0];
                  v.ext[1] = 3;
                  v.ext[2] = 3;
                  v.str[0] = 9;
                  v.str[1] = 3;
                  v.str[2] = 1;
                  return cg::createVectorGradientG2P( v, m );
              } );
    g2p_pair( 3, 1, "VectorDivergence", [&]( double* o ) { return VectorDivergenceG2P{ o, m }; },
              [&]( DeviceView<double, 2>& o ) {
                  DeviceView<double, 1> v;
                  v.mem = o.mem;
                  v.ext[0] = o.ext[0];
                  v.str[0] = 1;
                  return cg::createVectorDivergenceG2P( v, m );
              } );
}

int main()
{
    run_order<1>();
    run_order<2>();
    run_order<3>();
    std::printf( "%d checks, %d failures\n", g_checks, g_failures );
    if ( g_failures == 0 )
        std::printf( "ALL PASSED\n" );
    return g_failures == 0 ? 0 : 1;
}
