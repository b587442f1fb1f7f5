/*
 * cabana_amd_grid.hpp: header-only C++ layer over the cabana_amd C ABI with
 * the call shapes of Cabana::Grid's interpolation interface
 * (grid/src/Cabana_Grid_Interpolation.hpp), so that
 *
 *     namespace cg = cabana_amd::grid;
 *     auto f = cg::createScalarValueP2G( mass, 1.0 );
 *     cg::p2g( f, positions, num_point, cg::Spline<2>(), halo, *array );
 *
 * reads like the Cabana code it replaces.  Arguments are duck-typed:
 *
 *   particle fields / points: a Cabana Slice (data(), size(), viewRank(),
 *       extent(d), stride(0), static vector_length) or a Kokkos-style View
 *       (data(), extent(d), stride(d), static rank) in device memory, with
 *       double or float values;
 *   grid arrays: a Cabana::Grid::Array over Node entities (view() returning
 *       the (i, j, k, dof) view, layout()->localGrid() giving haloCellWidth(),
 *       globalGrid().isPeriodic(d), globalGrid().globalMesh().lowCorner(d) /
 *       highCorner(d) / cellSize(d)) of a single-rank uniform mesh;
 *   halos: cabana_amd::grid::Halo from createHalo below (carries its width),
 *       or any other object, which is taken at the full halo cell width;
 *   execution spaces: anything; one with hip_stream() supplies the stream.
 *
 * Errors from the library become cabana_amd::grid::Error exceptions.
 */
#ifndef CABANA_AMD_GRID_HPP
#define CABANA_AMD_GRID_HPP

#include "cabana_amd.h"

#include <hip/hip_runtime_api.h>

#include <array>
#include <atomic>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

namespace cabana_amd
{
namespace grid
{

// ---------------------------------------------------------------------------
// Tags (Cabana_Grid_Types.hpp, Cabana_Grid_Splines.hpp, Cabana_Grid_Halo*.hpp)
// ---------------------------------------------------------------------------

template <int Order>
struct Spline
{
    static_assert( Order >= 0 && Order <= 6,
                   "cabana_amd implements Spline<0> to Spline<6>" );
    static constexpr int order = Order;
};

template <std::size_t NumSpaceDim = 3>
struct NodeHaloPattern
{
    static_assert( NumSpaceDim == 3, "cabana_amd meshes are 3-D" );
};

namespace ScatterReduce
{
struct Sum
{
};
} // namespace ScatterReduce

class Error : public std::runtime_error
{
  public:
    Error( cabana_amd_status s, const std::string& what )
        : std::runtime_error( what )
        , status( s )
    {
    }
    cabana_amd_status status;
};

// ---------------------------------------------------------------------------
// Context: one per (thread, device), created on first use.
// ---------------------------------------------------------------------------

namespace detail
{

struct ContextHolder
{
    int device = -1;
    cabana_amd_context* ctx = nullptr;
    ~ContextHolder()
    {
        if ( ctx )
            cabana_amd_context_destroy( ctx );
    }
};

inline cabana_amd_context* context()
{
    thread_local std::vector<std::unique_ptr<ContextHolder>> held;
    int dev = 0;
    if ( hipGetDevice( &dev ) != hipSuccess )
        throw Error( CABANA_AMD_ERROR_DEVICE, "cabana_amd: no HIP device" );
    for ( auto& h : held )
        if ( h->device == dev )
            return h->ctx;
    auto h = std::make_unique<ContextHolder>();
    h->device = dev;
    cabana_amd_status s = cabana_amd_context_create( dev, &h->ctx );
    if ( s != CABANA_AMD_OK )
        throw Error( s, cabana_amd_last_error( nullptr ) );
    held.push_back( std::move( h ) );
    return held.back()->ctx;
}

inline void check( cabana_amd_status s, cabana_amd_context* ctx )
{
    if ( s != CABANA_AMD_OK )
        throw Error( s, cabana_amd_last_error( ctx ) );
}

template <class T>
constexpr int32_t dtype_of()
{
    using U = std::remove_cv_t<T>;
    static_assert( std::is_same<U, double>::value || std::is_same<U, float>::value ||
                       std::is_same<U, int32_t>::value || std::is_same<U, int64_t>::value,
                   "cabana_amd fields hold double, float, int32 or int64" );
    return std::is_same<U, double>::value    ? CABANA_AMD_FLOAT64
           : std::is_same<U, float>::value   ? CABANA_AMD_FLOAT32
           : std::is_same<U, int32_t>::value ? CABANA_AMD_INT32
                                             : CABANA_AMD_INT64;
}

// A Cabana Slice exposes a static vector_length and viewRank().
template <class S, class = void>
struct is_slice : std::false_type
{
};
template <class S>
struct is_slice<S, std::void_t<decltype( S::vector_length ),
                               decltype( std::declval<const S&>().viewRank() )>>
    : std::true_type
{
};

// Execution spaces with a HIP stream (Kokkos::HIP::hip_stream()).
template <class E, class = void>
struct has_hip_stream : std::false_type
{
};
template <class E>
struct has_hip_stream<E, std::void_t<decltype( std::declval<const E&>().hip_stream() )>>
    : std::true_type
{
};

template <class E>
cabana_amd_stream stream_of( const E& exec )
{
    if constexpr ( has_hip_stream<E>::value )
        return reinterpret_cast<cabana_amd_stream>( exec.hip_stream() );
    else
    {
        (void)exec;
        return nullptr;
    }
}

struct DefaultExec
{
};

// Particle field descriptor of a Slice: element (p, c) at
// data[(p / VL) * stride(0) + p % VL + VL * c] (Cabana_Slice.hpp:119-190).
template <class S>
cabana_amd_particle_field slice_field( const S& s )
{
    using T = std::remove_pointer_t<decltype( s.data() )>;
    cabana_amd_particle_field f{};
    f.data = const_cast<std::remove_cv_t<T>*>( s.data() );
    f.num_particle = static_cast<int64_t>( s.size() );
    int64_t nc = 1;
    for ( std::size_t d = 2; d < static_cast<std::size_t>( s.viewRank() ); ++d )
        nc *= static_cast<int64_t>( s.extent( d ) );
    f.num_comp = static_cast<int32_t>( nc );
    f.dtype = dtype_of<T>();
    f.vector_length = S::vector_length;
    f.soa_stride = static_cast<int64_t>( s.stride( 0 ) );
    return f;
}

// Particle field descriptor of a View indexed (p[, d0[, d1]]).  Row-major
// component packing (LayoutRight) maps to vector_length 1; particle-fastest
// storage (LayoutLeft) maps to vector_length = stride(1) when the
// components are in (d0, d1) row-major order.
template <class V>
cabana_amd_particle_field view_field( const V& v )
{
    using T = std::remove_pointer_t<decltype( v.data() )>;
    constexpr int rank = static_cast<int>( V::rank );
    static_assert( rank >= 1 && rank <= 3, "particle views are rank 1, 2 or 3" );
    cabana_amd_particle_field f{};
    f.data = const_cast<std::remove_cv_t<T>*>( v.data() );
    f.num_particle = static_cast<int64_t>( v.extent( 0 ) );
    f.dtype = dtype_of<T>();
    int64_t ext[3] = { 1, 1, 1 }, str[3] = { 0, 0, 0 };
    for ( int d = 0; d < rank; ++d )
    {
        ext[d] = static_cast<int64_t>( v.extent( d ) );
        str[d] = static_cast<int64_t>( v.stride( d ) );
    }
    const int64_t d1 = ( rank == 3 ) ? ext[2] : 1;
    f.num_comp = static_cast<int32_t>( ( rank >= 2 ? ext[1] : 1 ) * d1 );
    const bool comps_packed =
        ( rank == 1 ) || ( rank == 2 && ( ext[1] == 1 || str[1] == 1 ) ) ||
        ( rank == 3 && ( ext[2] == 1 || str[2] == 1 ) && ( ext[1] == 1 || str[1] == d1 ) );
    if ( comps_packed )
    {
        // Element (p, c) at p * stride(0) + c.
        f.vector_length = 1;
        f.soa_stride = str[0];
        return f;
    }
    // Particle-fastest: offset = p + VL * c with VL = stride of the last
    // component index.
    const int64_t vl = ( rank == 3 ) ? str[2] : str[1];
    const bool ok = str[0] == 1 && vl >= ext[0] &&
                    ( rank == 2 || ( ext[1] == 1 || str[1] == vl * d1 ) );
    if ( !ok )
        throw Error( CABANA_AMD_ERROR_INVALID_ARGUMENT,
                     "cabana_amd: particle view strides do not match "
                     "data[(p/VL)*stride + p%VL + VL*(d0*D1 + d1)]" );
    f.vector_length = vl;
    f.soa_stride = vl * f.num_comp;
    return f;
}

template <class X>
cabana_amd_particle_field particle_field( const X& x )
{
    if constexpr ( is_slice<X>::value )
        return slice_field( x );
    else
        return view_field( x );
}

template <class ArrayT>
cabana_amd_grid_field grid_field( const ArrayT& array )
{
    auto v = array.view();
    using T = std::remove_pointer_t<decltype( v.data() )>;
    cabana_amd_grid_field g{};
    g.data = const_cast<std::remove_cv_t<T>*>( v.data() );
    for ( int d = 0; d < 4; ++d )
    {
        g.extent[d] = static_cast<int32_t>( v.extent( d ) );
        g.stride[d] = static_cast<int64_t>( v.stride( d ) );
    }
    g.dtype = dtype_of<T>();
    return g;
}

template <class ArrayT>
cabana_amd_mesh mesh_of( const ArrayT& array )
{
    const auto& lg = *array.layout()->localGrid();
    const auto& gg = lg.globalGrid();
    const auto& gm = gg.globalMesh();
    cabana_amd_mesh m{};
    for ( int d = 0; d < 3; ++d )
    {
        m.low_corner[d] = gm.lowCorner( d );
        m.high_corner[d] = gm.highCorner( d );
        m.cell_size[d] = gm.cellSize( d );
        m.periodic[d] = gg.isPeriodic( d ) ? 1 : 0;
    }
    m.halo_cell_width = lg.haloCellWidth();
    return m;
}

template <class H, class = void>
struct has_width : std::false_type
{
};
template <class H>
struct has_width<H, std::void_t<decltype( std::declval<const H&>().width() )>>
    : std::true_type
{
};

template <class H>
int32_t halo_width_of( const H& halo )
{
    if constexpr ( has_width<H>::value )
        return static_cast<int32_t>( halo.width() );
    else
    {
        (void)halo;
        return -1;
    }
}

} // namespace detail

// ---------------------------------------------------------------------------
// Halo (single rank, periodic neighbours are the rank itself)
// ---------------------------------------------------------------------------

class Halo
{
  public:
    explicit Halo( int width )
        : _width( width )
    {
    }
    int width() const { return _width; }

    // Halo::gather (Cabana_Grid_Halo_Mpi.hpp:76-157): ghost = periodic owner.
    template <class ExecSpace, class ArrayT>
    void gather( const ExecSpace& exec, const ArrayT& array ) const
    {
        auto m = detail::mesh_of( array );
        auto g = detail::grid_field( array );
        auto* ctx = detail::context();
        detail::check( cabana_amd_halo_gather( ctx, &m, &g, _width, detail::stream_of( exec ) ),
                       ctx );
    }

    // Halo::scatter(Sum) (Cabana_Grid_Halo_Mpi.hpp:166-247).
    template <class ExecSpace, class ArrayT>
    void scatter( const ExecSpace& exec, ScatterReduce::Sum, const ArrayT& array ) const
    {
        auto m = detail::mesh_of( array );
        auto g = detail::grid_field( array );
        auto* ctx = detail::context();
        detail::check(
            cabana_amd_halo_scatter_sum( ctx, &m, &g, _width, detail::stream_of( exec ) ), ctx );
    }

  private:
    int _width;
};

template <class Pattern, class... Arrays>
std::shared_ptr<Halo> createHalo( const Pattern&, const int width, const Arrays&... )
{
    return std::make_shared<Halo>( width );
}

// ---------------------------------------------------------------------------
// Functors (Cabana_Grid_Interpolation.hpp:938-1566): the particle field, the
// operator and the multiplier.
// ---------------------------------------------------------------------------

struct P2GFunctor
{
    cabana_amd_p2g_op op;
    double multiplier;
    cabana_amd_particle_field values;
};

struct G2PFunctor
{
    cabana_amd_g2p_op op;
    double multiplier;
    cabana_amd_particle_field values;
};

namespace detail
{
template <class X>
cabana_amd_particle_field checked_field( const X& x, int32_t num_comp, const char* who )
{
    auto f = particle_field( x );
    if ( f.num_comp != num_comp )
        throw Error( CABANA_AMD_ERROR_INVALID_ARGUMENT,
                     std::string( who ) + ": particle field has the wrong rank/extent" );
    return f;
}
} // namespace detail

template <class X>
P2GFunctor createScalarValueP2G( const X& x, double multiplier )
{
    return { CABANA_AMD_SCALAR_VALUE_P2G, multiplier,
             detail::checked_field( x, 1, "createScalarValueP2G" ) };
}
template <class X>
P2GFunctor createVectorValueP2G( const X& x, double multiplier )
{
    return { CABANA_AMD_VECTOR_VALUE_P2G, multiplier,
             detail::checked_field( x, 3, "createVectorValueP2G" ) };
}
template <class X>
P2GFunctor createScalarGradientP2G( const X& x, double multiplier )
{
    return { CABANA_AMD_SCALAR_GRADIENT_P2G, multiplier,
             detail::checked_field( x, 1, "createScalarGradientP2G" ) };
}
template <class X>
P2GFunctor createVectorDivergenceP2G( const X& x, double multiplier )
{
    return { CABANA_AMD_VECTOR_DIVERGENCE_P2G, multiplier,
             detail::checked_field( x, 3, "createVectorDivergenceP2G" ) };
}
template <class X>
P2GFunctor createTensorDivergenceP2G( const X& x, double multiplier )
{
    return { CABANA_AMD_TENSOR_DIVERGENCE_P2G, multiplier,
             detail::checked_field( x, 9, "createTensorDivergenceP2G" ) };
}
template <class X>
G2PFunctor createScalarValueG2P( const X& x, double multiplier )
{
    return { CABANA_AMD_SCALAR_VALUE_G2P, multiplier,
             detail::checked_field( x, 1, "createScalarValueG2P" ) };
}
template <class X>
G2PFunctor createVectorValueG2P( const X& x, double multiplier )
{
    return { CABANA_AMD_VECTOR_VALUE_G2P, multiplier,
             detail::checked_field( x, 3, "createVectorValueG2P" ) };
}
template <class X>
G2PFunctor createScalarGradientG2P( const X& x, double multiplier )
{
    return { CABANA_AMD_SCALAR_GRADIENT_G2P, multiplier,
             detail::checked_field( x, 3, "createScalarGradientG2P" ) };
}
template <class X>
G2PFunctor createVectorGradientG2P( const X& x, double multiplier )
{
    return { CABANA_AMD_VECTOR_GRADIENT_G2P, multiplier,
             detail::checked_field( x, 9, "createVectorGradientG2P" ) };
}
template <class X>
G2PFunctor createVectorDivergenceG2P( const X& x, double multiplier )
{
    return { CABANA_AMD_VECTOR_DIVERGENCE_G2P, multiplier,
             detail::checked_field( x, 1, "createVectorDivergenceG2P" ) };
}

// ---------------------------------------------------------------------------
// LinkedCellList (core/src/Cabana_LinkedCellList.hpp:651-739) and permute
// (:1130-1145).  The list owns its device buffers.
// ---------------------------------------------------------------------------

class LinkedCellList
{
  public:
    template <class Positions>
    LinkedCellList( const Positions& positions, const std::array<double, 3>& grid_delta,
                    const std::array<double, 3>& grid_min,
                    const std::array<double, 3>& grid_max )
        : _delta( grid_delta )
        , _min( grid_min )
        , _max( grid_max )
    {
        detail::check( cabana_amd_linked_cell_num_bins( _delta.data(), _min.data(),
                                                        _max.data(), _nbin.data() ),
                       nullptr );
        const std::size_t nc = std::size_t( _nbin[0] ) * _nbin[1] * _nbin[2];
        alloc( &_counts, nc * sizeof( int32_t ) );
        alloc( &_offsets, nc * sizeof( int64_t ) );
        auto f = detail::particle_field( positions );
        build( positions, 0, f.num_particle );
    }
    LinkedCellList( const LinkedCellList& ) = delete;
    LinkedCellList& operator=( const LinkedCellList& ) = delete;
    ~LinkedCellList()
    {
        (void)hipFree( _counts );
        (void)hipFree( _offsets );
        (void)hipFree( _permutes );
    }

    template <class Positions>
    void build( const Positions& positions, const int64_t begin, const int64_t end )
    {
        if ( end - begin > _capacity )
        {
            (void)hipFree( _permutes );
            _permutes = nullptr;
            alloc( &_permutes, std::size_t( end - begin ) * sizeof( int64_t ) );
            _capacity = end - begin;
        }
        _begin = begin;
        _end = end;
        // A new serial for every build, taken before it (counts/offsets
        // change even when the build fails part way) from the library, so it
        // is unique in the process whatever the Positions type.  The library
        // reuses part tables derived from one build while the serial holds.
        _serial = cabana_amd_new_binning_serial();
        auto f = detail::particle_field( positions );
        auto* ctx = detail::context();
        detail::check( cabana_amd_linked_cell_build(
                           ctx, &f, begin, end, _delta.data(), _min.data(), _max.data(),
                           static_cast<int32_t*>( _counts ), static_cast<int64_t*>( _offsets ),
                           static_cast<int64_t*>( _permutes ), nullptr ),
                       ctx );
    }

    int numBin( const int d ) const { return _nbin[d]; }
    int totalBins() const { return _nbin[0] * _nbin[1] * _nbin[2]; }
    int cardinalIndex( const int i, const int j, const int k ) const
    {
        return ( i * _nbin[1] + j ) * _nbin[2] + k;
    }
    int binSize( const int i, const int j, const int k ) const
    {
        int32_t v = 0;
        copy_out( &v, static_cast<const int32_t*>( _counts ) + cardinalIndex( i, j, k ),
                  sizeof( v ) );
        return v;
    }
    int64_t binOffset( const int i, const int j, const int k ) const
    {
        int64_t v = 0;
        copy_out( &v, static_cast<const int64_t*>( _offsets ) + cardinalIndex( i, j, k ),
                  sizeof( v ) );
        return v;
    }
    int64_t permutation( const int64_t i ) const
    {
        int64_t v = 0;
        copy_out( &v, static_cast<const int64_t*>( _permutes ) + i, sizeof( v ) );
        return v;
    }
    int64_t rangeBegin() const { return _begin; }
    int64_t rangeEnd() const { return _end; }
    const int64_t* permutes() const { return static_cast<const int64_t*>( _permutes ); }

    // Binning handle for p2g / g2p once the particles are permuted.
    cabana_amd_binning binning() const
    {
        cabana_amd_binning b{};
        b.counts = static_cast<const int32_t*>( _counts );
        b.offsets = static_cast<const int64_t*>( _offsets );
        b.begin = _begin;
        b.end = _end;
        for ( int d = 0; d < 3; ++d )
            b.num_bin[d] = _nbin[d];
        b.serial = _serial;
        return b;
    }

  private:
    static void alloc( void** p, std::size_t bytes )
    {
        if ( hipMalloc( p, bytes ? bytes : 1 ) != hipSuccess )
            throw Error( CABANA_AMD_ERROR_OUT_OF_MEMORY, "cabana_amd: LinkedCellList hipMalloc" );
    }
    static void copy_out( void* dst, const void* src, std::size_t bytes )
    {
        if ( hipMemcpy( dst, src, bytes, hipMemcpyDeviceToHost ) != hipSuccess )
            throw Error( CABANA_AMD_ERROR_DEVICE, "cabana_amd: LinkedCellList read-back" );
    }

    std::array<double, 3> _delta, _min, _max;
    std::array<int32_t, 3> _nbin{};
    void* _counts = nullptr;
    void* _offsets = nullptr;
    void* _permutes = nullptr;
    int64_t _capacity = 0;
    int64_t _begin = 0;
    int64_t _end = 0;
    uint64_t _serial = 0;
};

template <class Positions>
std::shared_ptr<LinkedCellList> createLinkedCellList( const Positions& positions,
                                                      const std::array<double, 3>& grid_delta,
                                                      const std::array<double, 3>& grid_min,
                                                      const std::array<double, 3>& grid_max )
{
    return std::make_shared<LinkedCellList>( positions, grid_delta, grid_min, grid_max );
}

template <class Field>
void permute( const LinkedCellList& lcl, const Field& field )
{
    auto f = detail::particle_field( field );
    auto* ctx = detail::context();
    detail::check( cabana_amd_permute( ctx, lcl.permutes(), lcl.rangeBegin(), lcl.rangeEnd(),
                                       &f, nullptr ),
                   ctx );
}

// permute(binning_data, aosoa) (core/src/Cabana_Sort.hpp:549-589) over the
// members given: each is permuted once, the row-contiguous ones in one fused
// pass.  The members must not share memory.
template <class Field0, class Field1, class... Fields>
void permute( const LinkedCellList& lcl, const Field0& f0, const Field1& f1,
              const Fields&... rest )
{
    const cabana_amd_particle_field f[] = { detail::particle_field( f0 ),
                                            detail::particle_field( f1 ),
                                            detail::particle_field( rest )... };
    auto* ctx = detail::context();
    detail::check( cabana_amd_permute_members( ctx, lcl.permutes(), lcl.rangeBegin(),
                                               lcl.rangeEnd(), f,
                                               int32_t( 2 + sizeof...( rest ) ), nullptr ),
                   ctx );
}

// ---------------------------------------------------------------------------
// p2g / g2p (Cabana_Grid_Interpolation.hpp:831-927, 1217-1313)
// ---------------------------------------------------------------------------

// A functor bound to its grid array, for fused multi-functor calls.
struct P2GBinding
{
    P2GFunctor functor;
    cabana_amd_grid_field grid;
};
struct G2PBinding
{
    G2PFunctor functor;
    cabana_amd_grid_field grid;
};
template <class ArrayT>
P2GBinding bind( const P2GFunctor& f, const ArrayT& array )
{
    return { f, detail::grid_field( array ) };
}
template <class ArrayT>
G2PBinding bind( const G2PFunctor& f, const ArrayT& array )
{
    return { f, detail::grid_field( array ) };
}

template <class ExecSpace, class PointCoordinates, int Order, class HaloT, class ArrayT>
void p2g( const ExecSpace& exec_space, const P2GFunctor& functor,
          const PointCoordinates& points, const std::size_t num_point, Spline<Order>,
          const HaloT& halo, ArrayT& array, const LinkedCellList* bins = nullptr )
{
    auto m = detail::mesh_of( array );
    auto pts = detail::particle_field( points );
    pts.num_particle = static_cast<int64_t>( num_point );
    cabana_amd_p2g_term t{ functor.op, functor.multiplier, functor.values,
                           detail::grid_field( array ) };
    cabana_amd_binning b{};
    if ( bins )
        b = bins->binning();
    auto* ctx = detail::context();
    detail::check( cabana_amd_p2g( ctx, &m, Order, &pts, 1, &t, detail::halo_width_of( halo ),
                                   bins ? &b : nullptr, detail::stream_of( exec_space ) ),
                   ctx );
}

template <class PointCoordinates, int Order, class HaloT, class ArrayT>
void p2g( const P2GFunctor& functor, const PointCoordinates& points,
          const std::size_t num_point, Spline<Order> sp, const HaloT& halo, ArrayT& array,
          const LinkedCellList* bins = nullptr )
{
    p2g( detail::DefaultExec{}, functor, points, num_point, sp, halo, array, bins );
}

template <class ExecSpace, class ArrayT, class HaloT, class PointCoordinates, int Order>
void g2p( const ExecSpace& exec_space, const ArrayT& array, const HaloT& halo,
          const PointCoordinates& points, const std::size_t num_point, Spline<Order>,
          const G2PFunctor& functor, const LinkedCellList* bins = nullptr )
{
    auto m = detail::mesh_of( array );
    auto pts = detail::particle_field( points );
    pts.num_particle = static_cast<int64_t>( num_point );
    cabana_amd_g2p_term t{ functor.op, functor.multiplier, detail::grid_field( array ),
                           functor.values };
    cabana_amd_binning b{};
    if ( bins )
        b = bins->binning();
    auto* ctx = detail::context();
    detail::check( cabana_amd_g2p( ctx, &m, Order, &pts, 1, &t, detail::halo_width_of( halo ),
                                   bins ? &b : nullptr, detail::stream_of( exec_space ) ),
                   ctx );
}

template <class ArrayT, class HaloT, class PointCoordinates, int Order>
void g2p( const ArrayT& array, const HaloT& halo, const PointCoordinates& points,
          const std::size_t num_point, Spline<Order> sp, const G2PFunctor& functor,
          const LinkedCellList* bins = nullptr )
{
    g2p( detail::DefaultExec{}, array, halo, points, num_point, sp, functor, bins );
}

// Fused: several functors in one pass over the particles; `array` only
// supplies the mesh (every bound grid must live on it).
template <class ExecSpace, class PointCoordinates, int Order, class HaloT, class ArrayT>
void p2g( const ExecSpace& exec_space, const std::vector<P2GBinding>& terms,
          const PointCoordinates& points, const std::size_t num_point, Spline<Order>,
          const HaloT& halo, const ArrayT& mesh_array, const LinkedCellList* bins = nullptr )
{
    auto m = detail::mesh_of( mesh_array );
    auto pts = detail::particle_field( points );
    pts.num_particle = static_cast<int64_t>( num_point );
    std::vector<cabana_amd_p2g_term> t;
    for ( const auto& b : terms )
        t.push_back( { b.functor.op, b.functor.multiplier, b.functor.values, b.grid } );
    cabana_amd_binning bb{};
    if ( bins )
        bb = bins->binning();
    auto* ctx = detail::context();
    detail::check( cabana_amd_p2g( ctx, &m, Order, &pts, static_cast<int32_t>( t.size() ),
                                   t.data(), detail::halo_width_of( halo ),
                                   bins ? &bb : nullptr, detail::stream_of( exec_space ) ),
                   ctx );
}

template <class ExecSpace, class PointCoordinates, int Order, class HaloT, class ArrayT>
void g2p( const ExecSpace& exec_space, const std::vector<G2PBinding>& terms,
          const HaloT& halo, const PointCoordinates& points, const std::size_t num_point,
          Spline<Order>, const ArrayT& mesh_array, const LinkedCellList* bins = nullptr )
{
    auto m = detail::mesh_of( mesh_array );
    auto pts = detail::particle_field( points );
    pts.num_particle = static_cast<int64_t>( num_point );
    std::vector<cabana_amd_g2p_term> t;
    for ( const auto& b : terms )
        t.push_back( { b.functor.op, b.functor.multiplier, b.grid, b.functor.values } );
    cabana_amd_binning bb{};
    if ( bins )
        bb = bins->binning();
    auto* ctx = detail::context();
    detail::check( cabana_amd_g2p( ctx, &m, Order, &pts, static_cast<int32_t>( t.size() ),
                                   t.data(), detail::halo_width_of( halo ),
                                   bins ? &bb : nullptr, detail::stream_of( exec_space ) ),
                   ctx );
}

// ---------------------------------------------------------------------------
// Adjoints (no Cabana counterpart)
// ---------------------------------------------------------------------------

// values_bar += d<grid_bar, P2G(values)>/dvalues over owned nodes, and
// points_bar (FP64 (n, 3) view, optional) += the position derivative.
template <class PointCoordinates, int Order, class HaloT, class ArrayT, class ValuesBar,
          class PointsBar>
void p2g_adjoint( const P2GFunctor& functor, const PointCoordinates& points,
                  const std::size_t num_point, Spline<Order>, const HaloT& halo,
                  ArrayT& grid_bar, const ValuesBar& values_bar, const PointsBar* points_bar,
                  const LinkedCellList* bins = nullptr )
{
    auto m = detail::mesh_of( grid_bar );
    auto pts = detail::particle_field( points );
    pts.num_particle = static_cast<int64_t>( num_point );
    cabana_amd_p2g_adjoint_term t{ functor.op, functor.multiplier, functor.values,
                                   detail::grid_field( grid_bar ),
                                   detail::particle_field( values_bar ) };
    cabana_amd_particle_field xb{};
    if ( points_bar )
        xb = detail::particle_field( *points_bar );
    cabana_amd_binning b{};
    if ( bins )
        b = bins->binning();
    auto* ctx = detail::context();
    detail::check( cabana_amd_p2g_adjoint( ctx, &m, Order, &pts, 1, &t,
                                           detail::halo_width_of( halo ),
                                           points_bar ? &xb : nullptr, bins ? &b : nullptr,
                                           nullptr ),
                   ctx );
}

// grid_bar (owned nodes) += d<values_bar, G2P(grid)>/dgrid and points_bar
// (optional) += the position derivative; `functor` wraps values_bar.
template <class ArrayT, class HaloT, class PointCoordinates, int Order, class GridBar,
          class PointsBar>
void g2p_adjoint( const ArrayT& array, const HaloT& halo, const PointCoordinates& points,
                  const std::size_t num_point, Spline<Order>, const G2PFunctor& functor,
                  GridBar& grid_bar, const PointsBar* points_bar,
                  const LinkedCellList* bins = nullptr )
{
    auto m = detail::mesh_of( array );
    auto pts = detail::particle_field( points );
    pts.num_particle = static_cast<int64_t>( num_point );
    cabana_amd_g2p_adjoint_term t{ functor.op, functor.multiplier, detail::grid_field( array ),
                                   functor.values, detail::grid_field( grid_bar ) };
    cabana_amd_particle_field xb{};
    if ( points_bar )
        xb = detail::particle_field( *points_bar );
    cabana_amd_binning b{};
    if ( bins )
        b = bins->binning();
    auto* ctx = detail::context();
    detail::check( cabana_amd_g2p_adjoint( ctx, &m, Order, &pts, 1, &t,
                                           detail::halo_width_of( halo ),
                                           points_bar ? &xb : nullptr, bins ? &b : nullptr,
                                           nullptr ),
                   ctx );
}

// A P2G functor with its adjoint outputs, for adjoint_gather.
struct P2GAdjointBinding
{
    P2GFunctor functor;
    cabana_amd_grid_field grid_bar;
    cabana_amd_particle_field values_bar;
};
template <class ArrayT, class ValuesBar>
P2GAdjointBinding bind_adjoint( const P2GFunctor& f, const ArrayT& grid_bar,
                                const ValuesBar& values_bar )
{
    return { f, detail::grid_field( grid_bar ), detail::particle_field( values_bar ) };
}

// cabana_amd_adjoint_gather: p2g_adjoint of every P2G binding and the
// points_bar part of g2p_adjoint of every G2P functor (each wrapping its
// output cotangent and reading `array`), in one pass over the particles for
// the MPM set; points_bar (FP64 (n, 3)) += both position cotangents.
template <class ArrayT, class HaloT, class PointCoordinates, int Order, class PointsBar>
void adjoint_gather( const std::vector<P2GAdjointBinding>& p2g_terms, const ArrayT& array,
                     const HaloT& halo, const PointCoordinates& points,
                     const std::size_t num_point, Spline<Order>,
                     const std::vector<G2PFunctor>& g2p_functors, PointsBar& points_bar,
                     const LinkedCellList* bins = nullptr )
{
    auto m = detail::mesh_of( array );
    auto pts = detail::particle_field( points );
    pts.num_particle = static_cast<int64_t>( num_point );
    std::vector<cabana_amd_p2g_adjoint_term> pt;
    for ( const auto& b : p2g_terms )
        pt.push_back( { b.functor.op, b.functor.multiplier, b.functor.values, b.grid_bar,
                        b.values_bar } );
    std::vector<cabana_amd_g2p_adjoint_term> gt;
    for ( const auto& f : g2p_functors )
        gt.push_back( { f.op, f.multiplier, detail::grid_field( array ), f.values,
                        cabana_amd_grid_field{} } );
    cabana_amd_particle_field xb = detail::particle_field( points_bar );
    cabana_amd_binning b{};
    if ( bins )
        b = bins->binning();
    auto* ctx = detail::context();
    detail::check( cabana_amd_adjoint_gather(
                       ctx, &m, Order, &pts, static_cast<int32_t>( pt.size() ), pt.data(),
                       static_cast<int32_t>( gt.size() ), gt.data(), detail::halo_width_of( halo ),
                       &xb, bins ? &b : nullptr, nullptr ),
                   ctx );
}

// ---------------------------------------------------------------------------
// Library switches (no reference counterpart)
// ---------------------------------------------------------------------------

// Deterministic P2G for the calls that follow (cabana_amd_set_deterministic):
// every node sums its particles in a fixed order, so repeated calls give
// bitwise equal grids; Node arrays only, much slower than the default.
inline void setDeterministic( bool enable )
{
    auto* ctx = detail::context();
    detail::check( cabana_amd_set_deterministic( ctx, enable ? 1 : 0 ), ctx );
}

// Whether calls without a binning bin the particles themselves and run the
// cell-tiled kernels (cabana_amd_set_auto_binning; on by default).
inline void setAutoBinning( bool enable )
{
    auto* ctx = detail::context();
    detail::check( cabana_amd_set_auto_binning( ctx, enable ? 1 : 0 ), ctx );
}

} // namespace grid
} // namespace cabana_amd

#endif
