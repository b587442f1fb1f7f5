/*
 * cabana_amd_custom.hpp: user-written interpolation functors, the escape
 * hatch of Cabana::Grid's p2g / g2p (grid/src/Cabana_Grid_Interpolation.hpp
 * :831-927, :1217-1313), whose point loop calls any functor
 *
 *     functor( const SplineData& sd, const int p, const View& view )
 *
 * with the particle's spline data and the grid (a scatter view for P2G, the
 * array's view for G2P).  Here the point loop is a HIP kernel of this
 * header, so the translation unit that includes it is compiled with a HIP
 * compiler (hipcc) and the functor's operator() is __device__ (or __host__
 * __device__):
 *
 *     namespace cg = cabana_amd::grid;
 *     struct MyP2G {
 *         const double* m;
 *         template <class SD, class View>
 *         __device__ void operator()( const SD& sd, int p, const View& v ) const
 *         { cg::P2G::value( 2.0 * m[p], sd, v ); }
 *     };
 *     cg::p2g( MyP2G{ mass_ptr }, positions, num_point, cg::Spline<2>(), halo, *array );
 *
 * The P2G and G2P helpers below (value, gradient, divergence) are those of
 * Cabana_Grid_Interpolation.hpp :43-400 and :405-792 for 3-D, with the same
 * index conventions; a scatter view's access(i, j, k, d) += x adds
 * atomically, and the grid read by G2P has its ghosts filled first.  The
 * P2G call folds the ghosts into their owners afterwards
 * (Halo::scatter(Sum)), as Cabana's p2g does.  Particles whose stencil
 * leaves the ghosted grid are skipped (Cabana leaves that undefined).
 */
#ifndef CABANA_AMD_CUSTOM_HPP
#define CABANA_AMD_CUSTOM_HPP

#include "cabana_amd_grid.hpp"
#include "cabana_amd_spline.hpp"

#include <hip/hip_runtime.h>

#include <cmath>
#include <type_traits>
#include <vector>

namespace cabana_amd
{
namespace grid
{

// ---------------------------------------------------------------------------
// Spline data (Cabana_Grid_Splines.hpp:1051-1134, evaluateSpline :1280-1342)
// ---------------------------------------------------------------------------

template <class Scalar, int Order>
struct SplineData
{
    using scalar_type = Scalar;
    static constexpr int order = Order;
    static constexpr int num_knot = Order + 1;
    static constexpr int num_space_dim = 3;
    static constexpr bool has_weight_values = true;
    static constexpr bool has_weight_physical_gradients = true;
    Scalar x[3];            // logical position
    Scalar dx[3];           // cell size
    int s[3][num_knot];     // stencil node indices (local, ghosted)
    Scalar w[3][num_knot];  // weights
    Scalar g[3][num_knot];  // physical weight gradients
};

namespace detail
{
// The ghosted node grid of a single-rank uniform mesh.
struct NodeMesh
{
    double low[3]; // node (0, 0, 0) of the ghosted grid
    double rdx[3];
    double dx[3];
    int num_node[3];
};

inline NodeMesh node_mesh( const cabana_amd_mesh& m )
{
    NodeMesh n{};
    for ( int d = 0; d < 3; ++d )
    {
        // Cabana_Grid_GlobalMesh.hpp:59-80 (cell count), LocalMesh.hpp:61-120
        // (ghosted low corner) and Splines.hpp:1304-1318 (low corner = node
        // (0, 0, 0), dx = the difference of the first two node coordinates).
        const double c = m.cell_size[d];
        const int cells = int( std::rint( ( m.high_corner[d] - m.low_corner[d] ) / c ) );
        const double ghost_low = m.low_corner[d] - m.halo_cell_width * c;
        const double x1 = ghost_low + 1.0 * c;
        n.low[d] = ghost_low;
        n.dx[d] = x1 - ghost_low;
        n.rdx[d] = 1.0 / n.dx[d];
        n.num_node[d] = cells + 2 * m.halo_cell_width + 1;
    }
    return n;
}

// Element c of particle p of a slice-addressed field (Cabana_Slice.hpp
// :119-190): data[(p / vl) * soa_stride + p % vl + vl * c].
template <class T>
__device__ inline const T& field_at( const cabana_amd_particle_field& f, int64_t p, int c )
{
    const int64_t vl = f.vector_length;
    const T* base = static_cast<const T*>( f.data );
    return base[( p / vl ) * f.soa_stride + p % vl + vl * c];
}
} // namespace detail

// Spline data of position p (logical map, stencil, weights and gradients of
// every dimension); false when the stencil leaves the ghosted grid.
template <class Scalar, int Order>
__device__ inline bool evaluateSpline( const detail::NodeMesh& m, const double p[3],
                                       SplineData<Scalar, Order>& sd )
{
#pragma unroll
    for ( int d = 0; d < 3; ++d )
    {
        double x0;
        {
#pragma clang fp contract( off )
            x0 = ( p[d] - m.low[d] ) * m.rdx[d];
            if ( Order % 2 == 0 )
                x0 = x0 + 0.5;
        }
        if ( !( x0 >= 0.0 && x0 < double( m.num_node[d] ) ) )
            return false;
        const int i0 = int( x0 );
        const int base = i0 - Order / 2;
        if ( base < 0 || base + Order + 1 > m.num_node[d] )
            return false;
        sd.x[d] = Scalar( x0 );
        sd.dx[d] = Scalar( m.dx[d] );
#pragma unroll
        for ( int j = 0; j <= Order; ++j )
            sd.s[d][j] = base + j;
        spline_knot_weights<Order, Scalar>( Scalar( x0 - double( i0 ) ), Scalar( m.rdx[d] ),
                                            sd.w[d], sd.g[d] );
    }
    return true;
}

// ---------------------------------------------------------------------------
// Grid views passed to the functors
// ---------------------------------------------------------------------------

// G2P: the (i, j, k, dof) view of the array.
template <class T>
struct GridView
{
    const T* data;
    int64_t str[4];
    __device__ const T& operator()( int i, int j, int k, int d ) const
    {
        return data[i * str[0] + j * str[1] + k * str[2] + d * str[3]];
    }
};

// P2G: a scatter view whose access(i, j, k, d) += x adds atomically.
template <class T>
struct AtomicRef
{
    T* p;
    __device__ void operator+=( T x ) const { atomicAdd( p, x ); }
};
template <class T>
struct ScatterView
{
    T* data;
    int64_t str[4];
    __device__ const ScatterView& access() const { return *this; }
    __device__ AtomicRef<T> operator()( int i, int j, int k, int d ) const
    {
        return { data + ( i * str[0] + j * str[1] + k * str[2] + d * str[3] ) };
    }
};

// ---------------------------------------------------------------------------
// Helpers of Cabana_Grid_Interpolation.hpp (3-D)
// ---------------------------------------------------------------------------

namespace G2P
{
// Scalar value, :53-78.
template <class View, class SD, class R,
          std::enable_if_t<std::is_arithmetic<R>::value, int> = 0>
__device__ inline void value( const View& view, const SD& sd, R& result )
{
    result = R( 0 );
    for ( int i = 0; i < SD::num_knot; ++i )
        for ( int j = 0; j < SD::num_knot; ++j )
            for ( int k = 0; k < SD::num_knot; ++k )
                result += view( sd.s[0][i], sd.s[1][j], sd.s[2][k], 0 ) * sd.w[0][i] *
                          sd.w[1][j] * sd.w[2][k];
}
// Vector value, :108-136.
template <class View, class SD, class R>
__device__ inline void value( const View& view, const SD& sd, R result[3] )
{
    for ( int d = 0; d < 3; ++d )
        result[d] = R( 0 );
    for ( int i = 0; i < SD::num_knot; ++i )
        for ( int j = 0; j < SD::num_knot; ++j )
            for ( int k = 0; k < SD::num_knot; ++k )
                for ( int d = 0; d < 3; ++d )
                    result[d] += view( sd.s[0][i], sd.s[1][j], sd.s[2][k], d ) * sd.w[0][i] *
                                 sd.w[1][j] * sd.w[2][k];
}
// Scalar gradient, :168-203.
template <class View, class SD, class R>
__device__ inline void gradient( const View& view, const SD& sd, R result[3] )
{
    for ( int d = 0; d < 3; ++d )
        result[d] = R( 0 );
    for ( int i = 0; i < SD::num_knot; ++i )
        for ( int j = 0; j < SD::num_knot; ++j )
            for ( int k = 0; k < SD::num_knot; ++k )
            {
                const auto v = view( sd.s[0][i], sd.s[1][j], sd.s[2][k], 0 );
                result[0] += v * sd.g[0][i] * sd.w[1][j] * sd.w[2][k];
                result[1] += v * sd.w[0][i] * sd.g[1][j] * sd.w[2][k];
                result[2] += v * sd.w[0][i] * sd.w[1][j] * sd.g[2][k];
            }
}
// Vector gradient, :247-282: result[d0][d1] = d(view_d1) / dx_d0.
template <class View, class SD, class R>
__device__ inline void gradient( const View& view, const SD& sd, R result[3][3] )
{
    for ( int d0 = 0; d0 < 3; ++d0 )
        for ( int d1 = 0; d1 < 3; ++d1 )
            result[d0][d1] = R( 0 );
    for ( int i = 0; i < SD::num_knot; ++i )
        for ( int j = 0; j < SD::num_knot; ++j )
            for ( int k = 0; k < SD::num_knot; ++k )
            {
                const typename SD::scalar_type rg[3] = { sd.g[0][i] * sd.w[1][j] * sd.w[2][k],
                                                         sd.w[0][i] * sd.g[1][j] * sd.w[2][k],
                                                         sd.w[0][i] * sd.w[1][j] * sd.g[2][k] };
                for ( int d0 = 0; d0 < 3; ++d0 )
                    for ( int d1 = 0; d1 < 3; ++d1 )
                        result[d0][d1] += rg[d0] * view( sd.s[0][i], sd.s[1][j], sd.s[2][k], d1 );
            }
}
// Vector divergence, :333-364.
template <class View, class SD, class R,
          std::enable_if_t<std::is_arithmetic<R>::value, int> = 0>
__device__ inline void divergence( const View& view, const SD& sd, R& result )
{
    result = R( 0 );
    for ( int i = 0; i < SD::num_knot; ++i )
        for ( int j = 0; j < SD::num_knot; ++j )
            for ( int k = 0; k < SD::num_knot; ++k )
                result += view( sd.s[0][i], sd.s[1][j], sd.s[2][k], 0 ) * sd.g[0][i] *
                              sd.w[1][j] * sd.w[2][k] +
                          view( sd.s[0][i], sd.s[1][j], sd.s[2][k], 1 ) * sd.w[0][i] *
                              sd.g[1][j] * sd.w[2][k] +
                          view( sd.s[0][i], sd.s[1][j], sd.s[2][k], 2 ) * sd.w[0][i] *
                              sd.w[1][j] * sd.g[2][k];
}
} // namespace G2P

namespace P2G
{
// Scalar value, :437-459.
template <class P, class SD, class SV,
          std::enable_if_t<std::is_arithmetic<P>::value, int> = 0>
__device__ inline void value( const P& point_data, const SD& sd, const SV& view )
{
    const auto& a = view.access();
    for ( int i = 0; i < SD::num_knot; ++i )
        for ( int j = 0; j < SD::num_knot; ++j )
            for ( int k = 0; k < SD::num_knot; ++k )
                a( sd.s[0][i], sd.s[1][j], sd.s[2][k], 0 ) +=
                    point_data * sd.w[0][i] * sd.w[1][j] * sd.w[2][k];
}
// Vector value, :495-519.
template <class P, class SD, class SV>
__device__ inline void value( const P point_data[3], const SD& sd, const SV& view )
{
    const auto& a = view.access();
    for ( int i = 0; i < SD::num_knot; ++i )
        for ( int j = 0; j < SD::num_knot; ++j )
            for ( int k = 0; k < SD::num_knot; ++k )
                for ( int d = 0; d < 3; ++d )
                    a( sd.s[0][i], sd.s[1][j], sd.s[2][k], d ) +=
                        point_data[d] * sd.w[0][i] * sd.w[1][j] * sd.w[2][k];
}
// Scalar gradient, :557-589.
template <class P, class SD, class SV>
__device__ inline void gradient( const P point_data, const SD& sd, const SV& view )
{
    const auto& a = view.access();
    for ( int i = 0; i < SD::num_knot; ++i )
        for ( int j = 0; j < SD::num_knot; ++j )
            for ( int k = 0; k < SD::num_knot; ++k )
            {
                a( sd.s[0][i], sd.s[1][j], sd.s[2][k], 0 ) +=
                    point_data * sd.g[0][i] * sd.w[1][j] * sd.w[2][k];
                a( sd.s[0][i], sd.s[1][j], sd.s[2][k], 1 ) +=
                    point_data * sd.w[0][i] * sd.g[1][j] * sd.w[2][k];
                a( sd.s[0][i], sd.s[1][j], sd.s[2][k], 2 ) +=
                    point_data * sd.w[0][i] * sd.w[1][j] * sd.g[2][k];
            }
}
// Vector divergence, :637-665.
template <class P, class SD, class SV>
__device__ inline void divergence( const P point_data[3], const SD& sd, const SV& view )
{
    const auto& a = view.access();
    for ( int i = 0; i < SD::num_knot; ++i )
        for ( int j = 0; j < SD::num_knot; ++j )
            for ( int k = 0; k < SD::num_knot; ++k )
                a( sd.s[0][i], sd.s[1][j], sd.s[2][k], 0 ) +=
                    point_data[0] * sd.g[0][i] * sd.w[1][j] * sd.w[2][k] +
                    point_data[1] * sd.w[0][i] * sd.g[1][j] * sd.w[2][k] +
                    point_data[2] * sd.w[0][i] * sd.w[1][j] * sd.g[2][k];
}
// Tensor divergence, :713-748: view_d0 += sum_d1 point_data[d0][d1] d/dx_d1.
template <class P, class SD, class SV>
__device__ inline void divergence( const P point_data[3][3], const SD& sd, const SV& view )
{
    const auto& a = view.access();
    for ( int i = 0; i < SD::num_knot; ++i )
        for ( int j = 0; j < SD::num_knot; ++j )
            for ( int k = 0; k < SD::num_knot; ++k )
            {
                const typename SD::scalar_type rg[3] = { sd.g[0][i] * sd.w[1][j] * sd.w[2][k],
                                                         sd.w[0][i] * sd.g[1][j] * sd.w[2][k],
                                                         sd.w[0][i] * sd.w[1][j] * sd.g[2][k] };
                for ( int d0 = 0; d0 < 3; ++d0 )
                    a( sd.s[0][i], sd.s[1][j], sd.s[2][k], d0 ) +=
                        rg[0] * point_data[d0][0] + rg[1] * point_data[d0][1] +
                        rg[2] * point_data[d0][2];
            }
}
} // namespace P2G

// ---------------------------------------------------------------------------
// Point loops
// ---------------------------------------------------------------------------

namespace detail
{
// The library's own functor types and fused term lists, which keep the
// overloads of cabana_amd_grid.hpp.
template <class T>
struct is_builtin_functor
    : std::integral_constant<
          bool, std::is_same<std::decay_t<T>, P2GFunctor>::value ||
                    std::is_same<std::decay_t<T>, G2PFunctor>::value ||
                    std::is_same<std::decay_t<T>, std::vector<P2GBinding>>::value ||
                    std::is_same<std::decay_t<T>, std::vector<G2PBinding>>::value>
{
};

template <class Scalar, int Order, class Functor, class View>
__global__ void custom_point_kernel( Functor functor, cabana_amd_particle_field points,
                                     int64_t num_point, NodeMesh mesh, View view )
{
    const int64_t p = int64_t( blockIdx.x ) * blockDim.x + threadIdx.x;
    if ( p >= num_point )
        return;
    double x[3];
    for ( int d = 0; d < 3; ++d )
        x[d] = field_at<double>( points, p, d );
    SplineData<Scalar, Order> sd;
    if ( !evaluateSpline( mesh, x, sd ) )
        return;
    functor( sd, int( p ), view );
}

template <class ViewT, class ArrayT>
ViewT custom_view( const ArrayT& array )
{
    const cabana_amd_grid_field g = grid_field( array );
    ViewT v{};
    v.data = static_cast<decltype( v.data )>( g.data );
    for ( int d = 0; d < 4; ++d )
        v.str[d] = g.stride[d];
    return v;
}

template <class Scalar, int Order, class Functor, class View>
void launch_custom( const Functor& functor, const cabana_amd_particle_field& pts,
                    std::size_t num_point, const cabana_amd_mesh& m, const View& view,
                    cabana_amd_stream stream )
{
    if ( num_point == 0 )
        return;
    const unsigned block = 256;
    const unsigned grid = unsigned( ( num_point + block - 1 ) / block );
    hipLaunchKernelGGL( ( custom_point_kernel<Scalar, Order, Functor, View> ), dim3( grid ),
                        dim3( block ), 0, static_cast<hipStream_t>( stream ), functor, pts,
                        int64_t( num_point ), node_mesh( m ), view );
    if ( hipGetLastError() != hipSuccess )
        throw Error( CABANA_AMD_ERROR_DEVICE, "cabana_amd: custom functor launch failed" );
}
} // namespace detail

// p2g with a user functor (Cabana_Grid_Interpolation.hpp:1217-1313): the
// point loop adds into the array through a scatter view, then the ghosts
// are folded into their owners.
template <class ExecSpace, class Functor, class PointCoordinates, int Order, class HaloT,
          class ArrayT,
          std::enable_if_t<!detail::is_builtin_functor<Functor>::value, int> = 0>
void p2g( const ExecSpace& exec_space, const Functor& functor, const PointCoordinates& points,
          const std::size_t num_point, Spline<Order>, const HaloT& halo, ArrayT& array )
{
    using Scalar = std::remove_cv_t<std::remove_pointer_t<decltype( array.view().data() )>>;
    auto* ctx = detail::context();
    const cabana_amd_mesh m = detail::mesh_of( array );
    const cabana_amd_particle_field pts = detail::particle_field( points );
    const cabana_amd_stream s = detail::stream_of( exec_space );
    detail::launch_custom<Scalar, Order>( functor, pts, num_point, m,
                                          detail::custom_view<ScatterView<Scalar>>( array ), s );
    cabana_amd_grid_field g = detail::grid_field( array );
    detail::check( cabana_amd_halo_scatter_sum( ctx, &m, &g, detail::halo_width_of( halo ), s ),
                   ctx );
}

template <class Functor, class PointCoordinates, int Order, class HaloT, class ArrayT,
          std::enable_if_t<!detail::is_builtin_functor<Functor>::value, int> = 0>
void p2g( const Functor& functor, const PointCoordinates& points, const std::size_t num_point,
          Spline<Order> spline, const HaloT& halo, ArrayT& array )
{
    p2g( detail::DefaultExec{}, functor, points, num_point, spline, halo, array );
}

// g2p with a user functor (Cabana_Grid_Interpolation.hpp:831-927): the
// ghosts are filled from their owners, then the point loop reads the array.
template <class ExecSpace, class ArrayT, class HaloT, class PointCoordinates, int Order,
          class Functor,
          std::enable_if_t<!detail::is_builtin_functor<Functor>::value &&
                               !detail::is_builtin_functor<ArrayT>::value,
                           int> = 0>
void g2p( const ExecSpace& exec_space, const ArrayT& array, const HaloT& halo,
          const PointCoordinates& points, const std::size_t num_point, Spline<Order>,
          const Functor& functor )
{
    using Scalar = std::remove_cv_t<std::remove_pointer_t<decltype( array.view().data() )>>;
    auto* ctx = detail::context();
    const cabana_amd_mesh m = detail::mesh_of( array );
    const cabana_amd_stream s = detail::stream_of( exec_space );
    cabana_amd_grid_field g = detail::grid_field( array );
    detail::check( cabana_amd_halo_gather( ctx, &m, &g, detail::halo_width_of( halo ), s ), ctx );
    const cabana_amd_particle_field pts = detail::particle_field( points );
    detail::launch_custom<Scalar, Order>( functor, pts, num_point, m,
                                          detail::custom_view<GridView<Scalar>>( array ), s );
}

template <class ArrayT, class HaloT, class PointCoordinates, int Order, class Functor,
          std::enable_if_t<!detail::is_builtin_functor<Functor>::value &&
                               !detail::is_builtin_functor<ArrayT>::value,
                           int> = 0>
void g2p( const ArrayT& array, const HaloT& halo, const PointCoordinates& points,
          const std::size_t num_point, Spline<Order> spline, const Functor& functor )
{
    g2p( detail::DefaultExec{}, array, halo, points, num_point, spline, functor );
}

} // namespace grid
} // namespace cabana_amd

#endif
