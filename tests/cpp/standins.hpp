// Stand-ins for the Cabana / Kokkos types the C++ layer is duck-typed
// against (device Views, an AoSoA slice, Grid arrays over a single-rank
// uniform mesh) and the checking macros of the C++ test programs.
#pragma once

#include <hip/hip_runtime_api.h>

#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <vector>

static int g_failures = 0;
static int g_checks = 0;

#define EXPECT_NEAR( a, b, tol )                                                            \
    do                                                                                      \
    {                                                                                       \
        ++g_checks;                                                                         \
        const double va_ = ( a ), vb_ = ( b );                                              \
        if ( !( std::fabs( va_ - vb_ ) <= ( tol ) ) )                                        \
        {                                                                                   \
            if ( g_failures < 20 )                                                          \
                std::printf( "FAIL %s:%d: %s = %.17g, expected %.17g\n", __FILE__, __LINE__, \
                             #a, va_, vb_ );                                                \
            ++g_failures;                                                                   \
        }                                                                                   \
    } while ( 0 )

#define HIP_OK( call )                                                                      \
    do                                                                                      \
    {                                                                                       \
        hipError_t e_ = ( call );                                                           \
        if ( e_ != hipSuccess )                                                             \
        {                                                                                   \
            std::printf( "HIP error %s at %s:%d\n", hipGetErrorString( e_ ), __FILE__,       \
                         __LINE__ );                                                        \
            std::exit( 2 );                                                                 \
        }                                                                                   \
    } while ( 0 )

// ---------------------------------------------------------------------------
// Stand-ins
// ---------------------------------------------------------------------------

// Device view, LayoutRight, rank 1-4 (Kokkos::View interface subset).
template <class T, int Rank>
struct DeviceView
{
    static constexpr int rank = Rank;
    std::shared_ptr<T> mem;
    std::size_t ext[4] = { 1, 1, 1, 1 };
    std::size_t str[4] = { 0, 0, 0, 0 };

    DeviceView() = default;
    explicit DeviceView( std::array<std::size_t, Rank> e )
    {
        std::size_t n = 1;
        for ( int d = 0; d < Rank; ++d )
        {
            ext[d] = e[d];
            n *= e[d];
        }
        std::size_t s = 1;
        for ( int d = Rank - 1; d >= 0; --d )
        {
            str[d] = s;
            s *= ext[d];
        }
        T* p = nullptr;
        HIP_OK( hipMalloc( &p, n * sizeof( T ) + 1 ) );
        HIP_OK( hipMemset( p, 0, n * sizeof( T ) + 1 ) );
        mem = std::shared_ptr<T>( p, []( T* q ) { (void)hipFree( q ); } );
    }
    T* data() const { return mem.get(); }
    std::size_t extent( int d ) const { return ext[d]; }
    std::size_t stride( int d ) const { return str[d]; }
    std::size_t size() const { return ext[0] * ext[1] * ext[2] * ext[3]; }

    std::vector<T> to_host() const
    {
        std::vector<T> h( size() );
        HIP_OK( hipDeviceSynchronize() );
        HIP_OK( hipMemcpy( h.data(), data(), h.size() * sizeof( T ), hipMemcpyDeviceToHost ) );
        return h;
    }
    void from_host( const std::vector<T>& h )
    {
        HIP_OK( hipMemcpy( data(), h.data(), h.size() * sizeof( T ), hipMemcpyHostToDevice ) );
    }
    void fill( T v ) { from_host( std::vector<T>( size(), v ) ); }
};

// One member of an AoSoA of struct { double pos[3]; double mass; double
// vel[3]; } with vector length VL: the Cabana Slice interface subset.
template <int VL, int NComp>
struct AoSoASlice
{
    static constexpr int vector_length = VL;
    std::shared_ptr<double> mem;
    std::size_t offset;  // member offset within one SoA, in doubles
    std::size_t n;
    static constexpr std::size_t soa_doubles = 7 * VL;

    double* data() const { return mem.get() + offset; }
    std::size_t size() const { return n; }
    std::size_t viewRank() const { return NComp == 1 ? 2 : 3; }
    std::size_t extent( std::size_t d ) const
    {
        return d == 0 ? ( n + VL - 1 ) / VL : d == 1 ? VL : NComp;
    }
    std::size_t stride( std::size_t d ) const { return d == 0 ? soa_doubles : 1; }
};

struct GlobalMesh
{
    std::array<double, 3> lo, hi;
    double cell;
    double lowCorner( std::size_t d ) const { return lo[d]; }
    double highCorner( std::size_t d ) const { return hi[d]; }
    double cellSize( std::size_t ) const { return cell; }
};
struct GlobalGrid
{
    GlobalMesh mesh;
    const GlobalMesh& globalMesh() const { return mesh; }
    bool isPeriodic( int ) const { return true; }
};
struct LocalGrid
{
    GlobalGrid gg;
    int halo;
    const GlobalGrid& globalGrid() const { return gg; }
    int haloCellWidth() const { return halo; }
};
struct ArrayLayout
{
    std::shared_ptr<LocalGrid> lg;
    std::shared_ptr<LocalGrid> localGrid() const { return lg; }
};
template <class T>
struct Array
{
    std::shared_ptr<ArrayLayout> lay;
    DeviceView<T, 4> v;
    std::shared_ptr<ArrayLayout> layout() const { return lay; }
    const DeviceView<T, 4>& view() const { return v; }
};

template <class T>
std::shared_ptr<Array<T>> make_array( std::shared_ptr<LocalGrid> lg, int dofs, int n[3] )
{
    auto a = std::make_shared<Array<T>>();
    a->lay = std::make_shared<ArrayLayout>( ArrayLayout{ lg } );
    const int h = lg->halo;
    a->v = DeviceView<T, 4>( { std::size_t( n[0] + 2 * h + 1 ), std::size_t( n[1] + 2 * h + 1 ),
                               std::size_t( n[2] + 2 * h + 1 ), std::size_t( dofs ) } );
    return a;
}

