/*
 * cabana_amd_spline.hpp: the B-spline weight polynomials of Spline<0> to
 * Spline<6> (grid/src/Cabana_Grid_Splines.hpp) as compile-time tables,
 * shared by the library's kernels and by the header-only custom-functor
 * path (cabana_amd_custom.hpp).  Host and device; needs a HIP compiler.
 */
#ifndef CABANA_AMD_SPLINE_HPP
#define CABANA_AMD_SPLINE_HPP

#include <hip/hip_runtime.h>

namespace cabana_amd
{

// The 1-D B-spline weight of stencil node j as a polynomial of the
// fractional offset f: w_j(f) = sum_n spline_coef(ORDER, j, n) t^n with
// t = f - spline_shift(ORDER, j).  These are the weights of Spline<0> ..
// Spline<6>::value of Cabana_Grid_Splines.hpp (:94-100, :184-190, :281-299,
// :394-428, :522-557, :668-708, :833-881), written there as polynomials in
// x0 - int(x0) (minus 0.5 for orders 4 to 6) and here re-expanded around the
// end of [0, 1) where the node's weight is smallest (the middle for the
// centre node of an even order): the edge nodes' weights, (1 - f)^N / N! and
// f^N / N!, become a single power of t, and a small weight keeps its
// relative accuracy (a sum of O(1) terms would leave an absolute error of
// one rounding of 1 in FP32, enough to change a nearly empty node's mass by
// O(1)).
__host__ __device__ constexpr double spline_shift( int order, int j )
{
    return 2 * j < order ? 1.0 : ( 2 * j == order ? 0.5 : 0.0 );
}

// Coefficient of t^n in w_j.  Node j of a degree-N spline sees the
// polynomial piece w_j(f) = 1/N! sum_{k=0}^{N-j} (-1)^k C(N+1, k)
// (f + N - j - k)^N of the uniform B-spline (truncated-power form); with
// f = t + s each power expands binomially.  Every partial sum is a dyadic
// rational with fewer than 53 significant bits, so the sum is exact and
// the one division by N! rounds correctly.  Evaluated at compile time only
// (every caller binds it to a constexpr).
__host__ __device__ constexpr double spline_coef( int order, int j, int n )
{
    const double s = spline_shift( order, j );
    double fact = 1.0;
    for ( int i = 2; i <= order; ++i )
        fact *= i;
    double binom_n = 1.0; // C(order, n)
    for ( int i = 0; i < n; ++i )
        binom_n = binom_n * ( order - i ) / ( i + 1 );
    double sum = 0.0;
    double binom_k = 1.0; // C(order + 1, k)
    for ( int k = 0; k <= order - j; ++k )
    {
        const double c = s + ( order - j - k );
        double p = 1.0;
        for ( int e = 0; e < order - n; ++e )
            p *= c;
        sum += ( k % 2 ? -1.0 : 1.0 ) * binom_k * binom_n * p;
        binom_k = binom_k * ( order + 1 - k ) / ( k + 1 );
    }
    return sum / fact;
}

// Weight-polynomial coefficients and shifts of every stencil node of a
// spline order, as a compile-time table: c[j][n] is the coefficient of t^n
// in node j's weight, t = f - shift[j].
template <int Order>
struct SplineTable
{
    double c[Order + 1][Order + 1];
    double shift[Order + 1];
};

template <int Order>
__host__ __device__ constexpr SplineTable<Order> spline_table()
{
    SplineTable<Order> t{};
    for ( int j = 0; j <= Order; ++j )
    {
        t.shift[j] = spline_shift( Order, j );
        for ( int n = 0; n <= Order; ++n )
            t.c[j][n] = spline_coef( Order, j, n );
    }
    return t;
}

// Weights and physical gradients of the Order + 1 stencil nodes at
// fractional offset f (Horner's rule on the table; rdx = 1 / cell size).
template <int Order, class Real>
__host__ __device__ inline void spline_knot_weights( Real f, Real rdx, Real* w, Real* g )
{
    constexpr SplineTable<Order> T = spline_table<Order>();
#pragma unroll
    for ( int j = 0; j <= Order; ++j )
    {
        const Real t = f - Real( T.shift[j] );
        Real v = Real( T.c[j][Order] );
        Real d = Real( Order * T.c[j][Order] );
#pragma unroll
        for ( int n = Order - 1; n >= 0; --n )
        {
            v = v * t + Real( T.c[j][n] );
            if ( n >= 1 )
                d = d * t + Real( n * T.c[j][n] );
        }
        w[j] = v;
        g[j] = Order == 0 ? Real( 0 ) : d * rdx;
    }
}

} // namespace cabana_amd

#endif
