// Device-side building blocks: register-resident B-spline weights and the
// per-node bilinear forms shared by every scatter and gather kernel.
#pragma once

#include "ca_internal.hpp"

namespace cabana_amd
{

// Logical coordinate of a position, Spline<N>::mapToLogicalGrid
// (Cabana_Grid_Splines.hpp :141-145 primal for odd orders, :239-244 dual for
// even ones).  Rounded exactly as written: the cell index derived from it
// must not depend on FMA contraction.
template <int ORDER>
__device__ __forceinline__ double logical_coord( double xp, double low,
                                                 double rdx )
{
#pragma clang fp contract( off )
    double x = ( xp - low ) * rdx;
    if ( ORDER == 0 || ORDER == 2 )
        x = x + 0.5;
    return x;
}

// 1-D spline of one particle along one dimension: first stencil node, weights,
// physical gradients (dW/dx) and second derivatives (d2W/dx2).
template <int ORDER, class Real>
struct Spline1D
{
    static constexpr int K = ORDER + 1;
    int base;
    Real w[K];
    Real g[K];
    Real h[K];
};

// Evaluates Spline<ORDER>::stencil/value/gradient (Cabana_Grid_Splines.hpp
// <1> :163-213, <2> :264-324, <3> :383-451) from the FP64 logical coordinate.
// The fractional offset x0 - int(x0) is formed in FP64 and only then rounded
// to Real, so the FP32 path keeps FP64 cell indices and offsets.  Returns false
// when the stencil leaves [0, num_node) or the coordinate is not finite.
template <int ORDER, class Real, bool NEED_H>
__device__ __forceinline__ bool eval_spline_1d( double x0, Real rdx,
                                                int num_node,
                                                Spline1D<ORDER, Real>& s )
{
    constexpr int K = ORDER + 1;
    // Reject NaN / out-of-range before the int conversion.
    if ( !( x0 >= 0.0 && x0 < (double)num_node ) )
        return false;
    const int i0 = (int)x0;
    s.base = ( ORDER >= 2 ) ? i0 - 1 : i0;
    if ( s.base < 0 || s.base + K > num_node )
        return false;
    const Real f = (Real)( x0 - (double)i0 );
    if constexpr ( ORDER == 1 )
    {
        s.w[0] = Real( 1 ) - f;
        s.w[1] = f;
        s.g[0] = -rdx;
        s.g[1] = rdx;
        if constexpr ( NEED_H )
        {
            s.h[0] = Real( 0 );
            s.h[1] = Real( 0 );
        }
    }
    else if constexpr ( ORDER == 2 )
    {
        const Real nine_eights = Real( 9.0 / 8.0 );
        Real xn = f + Real( 0.5 );
        s.w[0] = Real( 0.5 ) * xn * xn - Real( 1.5 ) * xn + nine_eights;
        s.g[0] = ( xn - Real( 1.5 ) ) * rdx;
        xn -= Real( 1 );
        s.w[1] = -xn * xn + Real( 0.75 );
        s.g[1] = ( Real( -2 ) * xn ) * rdx;
        xn -= Real( 1 );
        s.w[2] = Real( 0.5 ) * xn * xn + Real( 1.5 ) * xn + nine_eights;
        s.g[2] = ( xn + Real( 1.5 ) ) * rdx;
        if constexpr ( NEED_H )
        {
            const Real r2 = rdx * rdx;
            s.h[0] = r2;
            s.h[1] = Real( -2 ) * r2;
            s.h[2] = r2;
        }
    }
    else
    {
        const Real one_sixth = Real( 1.0 / 6.0 );
        const Real two_thirds = one_sixth * Real( 4 );
        const Real four_thirds = two_thirds * Real( 2 );
        Real xn = f + Real( 1 );
        Real xn2 = xn * xn;
        s.w[0] = -xn * xn2 * one_sixth + xn2 - Real( 2 ) * xn + four_thirds;
        s.g[0] = ( Real( -0.5 ) * xn * xn + Real( 2 ) * xn - Real( 2 ) ) * rdx;
        Real h0 = -xn + Real( 2 );
        xn -= Real( 1 );
        xn2 = xn * xn;
        s.w[1] = Real( 0.5 ) * xn * xn2 - xn2 + two_thirds;
        s.g[1] = ( Real( 1.5 ) * xn * xn - Real( 2 ) * xn ) * rdx;
        Real h1 = Real( 3 ) * xn - Real( 2 );
        xn -= Real( 1 );
        xn2 = xn * xn;
        s.w[2] = Real( -0.5 ) * xn * xn2 - xn2 + two_thirds;
        s.g[2] = ( Real( -1.5 ) * xn * xn - Real( 2 ) * xn ) * rdx;
        Real h2 = Real( -3 ) * xn - Real( 2 );
        xn -= Real( 1 );
        xn2 = xn * xn;
        s.w[3] = xn * xn2 * one_sixth + xn2 + Real( 2 ) * xn + four_thirds;
        s.g[3] = ( Real( 0.5 ) * xn * xn + Real( 2 ) * xn + Real( 2 ) ) * rdx;
        Real h3 = xn + Real( 2 );
        if constexpr ( NEED_H )
        {
            const Real r2 = rdx * rdx;
            s.h[0] = h0 * r2;
            s.h[1] = h1 * r2;
            s.h[2] = h2 * r2;
            s.h[3] = h3 * r2;
        }
    }
    return true;
}

template <int ORDER, class Real, bool NEED_H>
struct Spline3D
{
    Spline1D<ORDER, Real> d[3];
};

// Evaluates the three 1-D splines of a particle position.
template <int ORDER, class Real, bool NEED_H>
__device__ __forceinline__ bool eval_spline( const MeshDev& m,
                                             const double px[3],
                                             Spline3D<ORDER, Real, NEED_H>& s )
{
    bool ok = true;
#pragma unroll
    for ( int d = 0; d < 3; ++d )
    {
        double x0 = logical_coord<ORDER>( px[d], m.low[d], m.rdx[d] );
        ok = ok && eval_spline_1d<ORDER, Real, NEED_H>(
                       x0, (Real)m.rdx[d], m.num_node[d], s.d[d] );
    }
    return ok;
}

// Weight products of one stencil node.
template <class Real, bool NEED_DD>
struct NodeW
{
    Real W;
    Real dW[3];
    Real ddW[6]; // (00, 11, 22, 01, 02, 12)
};

template <int ORDER, class Real, bool NEED_H, bool NEED_DD>
__device__ __forceinline__ NodeW<Real, NEED_DD>
node_weights( const Spline3D<ORDER, Real, NEED_H>& s, int i, int j, int k )
{
    NodeW<Real, NEED_DD> n;
    const Real wx = s.d[0].w[i], wy = s.d[1].w[j], wz = s.d[2].w[k];
    const Real gx = s.d[0].g[i], gy = s.d[1].g[j], gz = s.d[2].g[k];
    const Real wyz = wy * wz;
    n.W = wx * wyz;
    n.dW[0] = gx * wyz;
    n.dW[1] = wx * gy * wz;
    n.dW[2] = wx * wy * gz;
    if constexpr ( NEED_DD )
    {
        const Real hx = s.d[0].h[i], hy = s.d[1].h[j], hz = s.d[2].h[k];
        n.ddW[0] = hx * wyz;
        n.ddW[1] = wx * hy * wz;
        n.ddW[2] = wx * wy * hz;
        n.ddW[3] = gx * gy * wz;
        n.ddW[4] = gx * wy * gz;
        n.ddW[5] = wx * gy * gz;
    }
    return n;
}

template <class Real, bool DD>
__device__ __forceinline__ Real dd( const NodeW<Real, DD>& n, int e, int f )
{
    if ( e == f )
        return n.ddW[e];
    const int a = e < f ? e : f, b = e < f ? f : e;
    return ( a == 0 ) ? ( b == 1 ? n.ddW[3] : n.ddW[4] ) : n.ddW[5];
}

template <int FORM>
struct FormTraits;
template <>
struct FormTraits<FORM_SV>
{
    static constexpr int NP = 1, NG = 1;
};
template <>
struct FormTraits<FORM_VV>
{
    static constexpr int NP = 3, NG = 3;
};
template <>
struct FormTraits<FORM_SG>
{
    static constexpr int NP = 1, NG = 3;
};
template <>
struct FormTraits<FORM_VD>
{
    static constexpr int NP = 3, NG = 1;
};
template <>
struct FormTraits<FORM_TD>
{
    static constexpr int NP = 9, NG = 3;
};
template <>
struct FormTraits<FORM_TDT>
{
    static constexpr int NP = 9, NG = 3;
};

// dB/dG: contribution of particle vector P (already scaled by the
// multiplier) to the node dofs.  Operation order follows the P2G point
// kernels (Cabana_Grid_Interpolation.hpp:437-458, 495-518, 557-588, 631-665,
// 712-747).
template <int FORM, class Real, bool DD>
__device__ __forceinline__ void form_scatter( const Real* P,
                                              const NodeW<Real, DD>& n,
                                              Real* G )
{
    if constexpr ( FORM == FORM_SV )
    {
        G[0] = P[0] * n.W;
    }
    else if constexpr ( FORM == FORM_VV )
    {
#pragma unroll
        for ( int d = 0; d < 3; ++d )
            G[d] = P[d] * n.W;
    }
    else if constexpr ( FORM == FORM_SG )
    {
#pragma unroll
        for ( int d = 0; d < 3; ++d )
            G[d] = P[0] * n.dW[d];
    }
    else if constexpr ( FORM == FORM_VD )
    {
        G[0] = P[0] * n.dW[0] + P[1] * n.dW[1] + P[2] * n.dW[2];
    }
    else if constexpr ( FORM == FORM_TD )
    {
#pragma unroll
        for ( int a = 0; a < 3; ++a )
            G[a] = n.dW[0] * P[3 * a] + n.dW[1] * P[3 * a + 1] +
                   n.dW[2] * P[3 * a + 2];
    }
    else
    {
#pragma unroll
        for ( int b = 0; b < 3; ++b )
            G[b] = n.dW[0] * P[b] + n.dW[1] * P[3 + b] + n.dW[2] * P[6 + b];
    }
}

// dB/dP accumulated over nodes (G2P point kernels, :53-72, 108-130, 168-203,
// 247-282, 333-364).
template <int FORM, class Real, bool DD>
__device__ __forceinline__ void form_gather( const Real* G,
                                             const NodeW<Real, DD>& n,
                                             Real* acc )
{
    if constexpr ( FORM == FORM_SV )
    {
        acc[0] += G[0] * n.W;
    }
    else if constexpr ( FORM == FORM_VV )
    {
#pragma unroll
        for ( int d = 0; d < 3; ++d )
            acc[d] += G[d] * n.W;
    }
    else if constexpr ( FORM == FORM_SG )
    {
        acc[0] += G[0] * n.dW[0] + G[1] * n.dW[1] + G[2] * n.dW[2];
    }
    else if constexpr ( FORM == FORM_VD )
    {
#pragma unroll
        for ( int d = 0; d < 3; ++d )
            acc[d] += G[0] * n.dW[d];
    }
    else if constexpr ( FORM == FORM_TD )
    {
#pragma unroll
        for ( int a = 0; a < 3; ++a )
#pragma unroll
            for ( int b = 0; b < 3; ++b )
                acc[3 * a + b] += n.dW[b] * G[a];
    }
    else
    {
#pragma unroll
        for ( int a = 0; a < 3; ++a )
#pragma unroll
            for ( int b = 0; b < 3; ++b )
                acc[3 * a + b] += n.dW[a] * G[b];
    }
}

// dB/dx_e at one node: W -> dW_e, dW_d -> d2W/dx_e dx_d.
template <int FORM, class Real>
__device__ __forceinline__ void form_dx( const Real* P, const Real* G,
                                         const NodeW<Real, true>& n,
                                         Real* xb )
{
#pragma unroll
    for ( int e = 0; e < 3; ++e )
    {
        Real v;
        if constexpr ( FORM == FORM_SV )
        {
            v = P[0] * G[0] * n.dW[e];
        }
        else if constexpr ( FORM == FORM_VV )
        {
            v = ( P[0] * G[0] + P[1] * G[1] + P[2] * G[2] ) * n.dW[e];
        }
        else if constexpr ( FORM == FORM_SG )
        {
            v = P[0] * ( G[0] * dd( n, e, 0 ) + G[1] * dd( n, e, 1 ) +
                         G[2] * dd( n, e, 2 ) );
        }
        else if constexpr ( FORM == FORM_VD )
        {
            v = G[0] * ( P[0] * dd( n, e, 0 ) + P[1] * dd( n, e, 1 ) +
                         P[2] * dd( n, e, 2 ) );
        }
        else if constexpr ( FORM == FORM_TD )
        {
            v = Real( 0 );
#pragma unroll
            for ( int b = 0; b < 3; ++b )
                v += ( P[b] * G[0] + P[3 + b] * G[1] + P[6 + b] * G[2] ) *
                     dd( n, e, b );
        }
        else
        {
            v = Real( 0 );
#pragma unroll
            for ( int a = 0; a < 3; ++a )
                v += ( P[3 * a] * G[0] + P[3 * a + 1] * G[1] +
                       P[3 * a + 2] * G[2] ) *
                     dd( n, e, a );
        }
        xb[e] += v;
    }
}

template <class Real>
__device__ __forceinline__ void atomic_add( Real* p, Real v )
{
    unsafeAtomicAdd( p, v );
}

__device__ __forceinline__ void load_position( const PFieldDev& f, int64_t p,
                                               double px[3] )
{
    const int64_t b = f.base( p );
    px[0] = *f.ptr<double>( b, 0 );
    px[1] = *f.ptr<double>( b, 1 );
    px[2] = *f.ptr<double>( b, 2 );
}

} // namespace cabana_amd
