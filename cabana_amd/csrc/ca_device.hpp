// Device-side building blocks: register-resident B-spline weights and the
// per-node bilinear forms shared by every scatter and gather kernel.
#pragma once

#include "ca_internal.hpp"
#include "../../include/cabana_amd_spline.hpp"

#include <utility>

namespace cabana_amd
{

// Calls f(std::integral_constant<int, I>) for I = 0 .. N-1.
template <class F, std::size_t... I>
__device__ __forceinline__ void static_for_impl( F&& f,
                                                 std::index_sequence<I...> )
{
    ( f( std::integral_constant<int, (int)I>{} ), ... );
}
template <int N, class F>
__device__ __forceinline__ void static_for( F&& f )
{
    static_for_impl( f, std::make_index_sequence<N>{} );
}

// Logical coordinate of a position, Spline<N>::mapToLogicalGrid
// (Cabana_Grid_Splines.hpp :146-149, :352-355, :622-625 primal for odd
// orders, :61-64, :241-244, :478-481, :785-788 dual for even ones).  Rounded exactly as written: the cell index derived from it
// must not depend on FMA contraction.
template <int ORDER>
__device__ __forceinline__ double logical_coord( double xp, double low,
                                                 double rdx )
{
#pragma clang fp contract( off )
    double x = ( xp - low ) * rdx;
    if ( ORDER % 2 == 0 )
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

// Stencil base node and fractional offset of a logical coordinate
// (Spline<ORDER>::stencil, Cabana_Grid_Splines.hpp <0> :80-84, <1> :169-175,
// <2> :265-272, <3> :377-385, <4> :504-512, <5> :649-659, <6> :813-824: the
// first node is int(x0) - ORDER / 2).  The offset x0 - int(x0) is formed in FP64 and only then
// rounded to Real, so the FP32 path keeps FP64 cell indices and offsets.
// Returns false when the stencil leaves [0, num_node) or the coordinate is
// not finite.
template <int ORDER, class Real>
__device__ __forceinline__ bool spline_frac( double x0, int num_node, int& base,
                                             Real& f )
{
    constexpr int K = ORDER + 1;
    // Reject NaN / out-of-range before the int conversion.
    if ( !( x0 >= 0.0 && x0 < (double)num_node ) )
        return false;
    const int i0 = (int)x0;
    base = i0 - ORDER / 2;
    if ( base < 0 || base + K > num_node )
        return false;
    f = (Real)( x0 - (double)i0 );
    return true;
}

// Value, first and second derivative (in f) of node j's weight polynomial
// at t = f - shift, by Horner's rule with compile-time coefficients.
template <int ORDER, int J, class Real>
__device__ __forceinline__ void spline_node_poly( Real t, Real& w, Real& d1, Real& d2 )
{
    constexpr double cN = spline_coef( ORDER, J, ORDER );
    w = Real( cN );
    static_for<ORDER>( [&]( auto I ) {
        constexpr int n = ORDER - 1 - I;
        constexpr double c = spline_coef( ORDER, J, n );
        w = w * t + Real( c );
    } );
    d1 = Real( ORDER * cN );
    if constexpr ( ORDER >= 2 )
        static_for<ORDER - 1>( [&]( auto I ) {
            constexpr int n = ORDER - 1 - I;
            constexpr double c = n * spline_coef( ORDER, J, n );
            d1 = d1 * t + Real( c );
        } );
    d2 = Real( 0 );
    if constexpr ( ORDER >= 2 )
    {
        d2 = Real( ORDER * ( ORDER - 1 ) * cN );
        if constexpr ( ORDER >= 3 )
            static_for<ORDER - 2>( [&]( auto I ) {
                constexpr int n = ORDER - 1 - I;
                constexpr double c = n * ( n - 1 ) * spline_coef( ORDER, J, n );
                d2 = d2 * t + Real( c );
            } );
    }
}

// Spline<ORDER>::value/gradient (Cabana_Grid_Splines.hpp <0> :94-117, <1>
// :184-213, <2> :281-324, <3> :394-451, <4> :522-597, <5> :668-760, <6>
// :833-938) and second derivatives from the fractional offset.
template <int ORDER, class Real, bool NEED_H>
__device__ __forceinline__ void spline_weights( Real f, Real rdx,
                                                Spline1D<ORDER, Real>& s )
{
    static_for<ORDER + 1>( [&]( auto J ) {
        constexpr int j = J;
        constexpr double sh = spline_shift( ORDER, j );
        const Real t = f - Real( sh );
        Real w, d1, d2;
        spline_node_poly<ORDER, j, Real>( t, w, d1, d2 );
        s.w[j] = w;
        s.g[j] = d1 * rdx;
        if constexpr ( NEED_H )
            s.h[j] = d2 * ( rdx * rdx );
    } );
}

// Stencil base, weights, gradients (and second derivatives) of one
// coordinate.
template <int ORDER, class Real, bool NEED_H>
__device__ __forceinline__ bool eval_spline_1d( double x0, Real rdx,
                                                int num_node,
                                                Spline1D<ORDER, Real>& s )
{
    Real f;
    if ( !spline_frac<ORDER, Real>( x0, num_node, s.base, f ) )
        return false;
    spline_weights<ORDER, Real, NEED_H>( f, rdx, s );
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

// Stencil base node of a position along every dimension (the first node of
// Spline<ORDER>::stencil), with the same FP64 rounding as eval_spline.
// Returns false where eval_spline would.
template <int ORDER>
__device__ __forceinline__ bool stencil_base( const MeshDev& m,
                                              const double px[3], int b[3] )
{
    constexpr int K = ORDER + 1;
    bool ok = true;
#pragma unroll
    for ( int d = 0; d < 3; ++d )
    {
        const double x0 = logical_coord<ORDER>( px[d], m.low[d], m.rdx[d] );
        const bool fin = x0 >= 0.0 && x0 < (double)m.num_node[d];
        const int i0 = fin ? (int)x0 : 0;
        b[d] = i0 - ORDER / 2;
        ok = ok && fin && b[d] >= 0 && b[d] + K <= m.num_node[d];
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

// ---------------------------------------------------------------------------
// Jets: derivatives of the interpolated grid field at a particle.
//
// For one grid component f on the K^3 stencil, the value, gradient and
// Hessian of sum_n W_n(x) f_n are tensor-product contractions of f with the
// 1-D weights w, gradients g and second derivatives h.  They are evaluated
// separably (contract k, then j, then i), which needs far fewer operations
// than forming each node's weight products.
// ---------------------------------------------------------------------------

template <class Real>
struct Jet
{
    Real v;
    Real g[3];
    Real h[6]; // xx, yy, zz, xy, xz, yz
};

// J[c] for NG grid components; Load(i, j, k, f) fills f[0..NG) of stencil
// node (i, j, k).  V, G, H select value, gradient and Hessian.
template <int ORDER, int NG, bool V, bool G, bool H, class Real, bool SH,
          class Load>
__device__ __forceinline__ void eval_jets( const Spline3D<ORDER, Real, SH>& s,
                                           Load&& load, Jet<Real>* J )
{
    constexpr int K = ORDER + 1;
    constexpr bool NG1 = G || H; // need first derivatives in y/z stages
#pragma unroll
    for ( int c = 0; c < NG; ++c )
    {
        J[c].v = Real( 0 );
#pragma unroll
        for ( int e = 0; e < 3; ++e )
            J[c].g[e] = Real( 0 );
#pragma unroll
        for ( int e = 0; e < 6; ++e )
            J[c].h[e] = Real( 0 );
    }
#pragma unroll
    for ( int i = 0; i < K; ++i )
    {
        Real Uww[NG], Ugw[NG], Uwg[NG], Uhw[NG], Uwh[NG], Ugg[NG];
#pragma unroll
        for ( int c = 0; c < NG; ++c )
            Uww[c] = Ugw[c] = Uwg[c] = Uhw[c] = Uwh[c] = Ugg[c] = Real( 0 );
#pragma unroll
        for ( int j = 0; j < K; ++j )
        {
            Real Tw[NG], Tg[NG], Th[NG];
#pragma unroll
            for ( int c = 0; c < NG; ++c )
                Tw[c] = Tg[c] = Th[c] = Real( 0 );
#pragma unroll
            for ( int k = 0; k < K; ++k )
            {
                Real f[NG];
                load( i, j, k, f );
                const Real wz = s.d[2].w[k];
#pragma unroll
                for ( int c = 0; c < NG; ++c )
                {
                    Tw[c] += wz * f[c];
                    if constexpr ( NG1 )
                        Tg[c] += s.d[2].g[k] * f[c];
                    if constexpr ( H )
                        Th[c] += s.d[2].h[k] * f[c];
                }
            }
            // One stencil row per scheduling region: the row's loads are
            // not hoisted into the previous row, which bounds the live
            // values to one row.
            __builtin_amdgcn_sched_barrier( 0 );
            const Real wy = s.d[1].w[j];
#pragma unroll
            for ( int c = 0; c < NG; ++c )
            {
                Uww[c] += wy * Tw[c];
                if constexpr ( NG1 )
                {
                    Ugw[c] += s.d[1].g[j] * Tw[c];
                    Uwg[c] += wy * Tg[c];
                }
                if constexpr ( H )
                {
                    Uhw[c] += s.d[1].h[j] * Tw[c];
                    Uwh[c] += wy * Th[c];
                    Ugg[c] += s.d[1].g[j] * Tg[c];
                }
            }
        }
        const Real wx = s.d[0].w[i], gx = s.d[0].g[i];
#pragma unroll
        for ( int c = 0; c < NG; ++c )
        {
            if constexpr ( V )
                J[c].v += wx * Uww[c];
            if constexpr ( G )
            {
                J[c].g[0] += gx * Uww[c];
                J[c].g[1] += wx * Ugw[c];
                J[c].g[2] += wx * Uwg[c];
            }
            if constexpr ( H )
            {
                J[c].h[0] += s.d[0].h[i] * Uww[c];
                J[c].h[1] += wx * Uhw[c];
                J[c].h[2] += wx * Uwh[c];
                J[c].h[3] += gx * Ugw[c];
                J[c].h[4] += gx * Uwg[c];
                J[c].h[5] += wx * Ugg[c];
            }
        }
        // Keep the next slab's loads from being hoisted above this one: the
        // fully unrolled stencil would otherwise hold every node in VGPRs.
        __builtin_amdgcn_sched_barrier( 0 );
    }
}

__device__ __forceinline__ constexpr int hidx( int e, int f )
{
    return e == f ? e
                  : ( ( e < f ? e : f ) == 0 ? ( ( e < f ? f : e ) == 1 ? 3 : 4 )
                                             : 5 );
}

// Jet levels a form needs from its grid: value-type forms (SV, VV) read the
// value, gradient-type forms read the gradient; the position adjoint adds one
// derivative order.
__device__ __forceinline__ constexpr bool form_value_type( int form )
{
    return form == FORM_SV || form == FORM_VV;
}

// dB/dP of form FORM from the jets of its grid (the G2P output / the P2G
// value cotangent before the multiplier).
template <int FORM, class Real>
__device__ __forceinline__ void jet_dP( const Jet<Real>* J, Real* out )
{
    if constexpr ( FORM == FORM_SV )
        out[0] = J[0].v;
    else if constexpr ( FORM == FORM_VV )
    {
#pragma unroll
        for ( int d = 0; d < 3; ++d )
            out[d] = J[d].v;
    }
    else if constexpr ( FORM == FORM_SG )
        out[0] = J[0].g[0] + J[1].g[1] + J[2].g[2];
    else if constexpr ( FORM == FORM_VD )
    {
#pragma unroll
        for ( int d = 0; d < 3; ++d )
            out[d] = J[0].g[d];
    }
    else if constexpr ( FORM == FORM_TD )
    {
#pragma unroll
        for ( int a = 0; a < 3; ++a )
#pragma unroll
            for ( int b = 0; b < 3; ++b )
                out[3 * a + b] = J[a].g[b];
    }
    else
    {
#pragma unroll
        for ( int a = 0; a < 3; ++a )
#pragma unroll
            for ( int b = 0; b < 3; ++b )
                out[3 * a + b] = J[b].g[a];
    }
}

// dB/dx_e of form FORM for particle vector P from the jets of its grid.
template <int FORM, class Real>
__device__ __forceinline__ void jet_dx( const Real* P, const Jet<Real>* J,
                                        Real* xb )
{
#pragma unroll
    for ( int e = 0; e < 3; ++e )
    {
        Real v;
        if constexpr ( FORM == FORM_SV )
            v = P[0] * J[0].g[e];
        else if constexpr ( FORM == FORM_VV )
            v = P[0] * J[0].g[e] + P[1] * J[1].g[e] + P[2] * J[2].g[e];
        else if constexpr ( FORM == FORM_SG )
            v = P[0] * ( J[0].h[hidx( e, 0 )] + J[1].h[hidx( e, 1 )] +
                         J[2].h[hidx( e, 2 )] );
        else if constexpr ( FORM == FORM_VD )
            v = P[0] * J[0].h[hidx( e, 0 )] + P[1] * J[0].h[hidx( e, 1 )] +
                P[2] * J[0].h[hidx( e, 2 )];
        else if constexpr ( FORM == FORM_TD )
        {
            v = Real( 0 );
#pragma unroll
            for ( int a = 0; a < 3; ++a )
#pragma unroll
                for ( int b = 0; b < 3; ++b )
                    v += P[3 * a + b] * J[a].h[hidx( e, b )];
        }
        else
        {
            v = Real( 0 );
#pragma unroll
            for ( int a = 0; a < 3; ++a )
#pragma unroll
                for ( int b = 0; b < 3; ++b )
                    v += P[3 * a + b] * J[b].h[hidx( e, a )];
        }
        xb[e] += v;
    }
}

// Gather pass modes: outputs only (G2P forward), outputs and position
// cotangent (P2G adjoint), position cotangent only (position part of the
// G2P adjoint, or a P2G adjoint without value cotangents).
enum GatherMode : int
{
    GATHER_OUT = 0,
    GATHER_OUT_XBAR = 1,
    GATHER_XBAR = 2
};
constexpr bool mode_xbar( int mode ) { return mode != GATHER_OUT; }
constexpr bool mode_out( int mode ) { return mode != GATHER_XBAR; }

// Flags on the entries of a gather pass's form list.  FORM_XONLY: the term
// adds only its position cotangent (the position part of a G2P adjoint run
// beside a P2G adjoint); its particle field is the contraction vector and
// it has no output.  FORM_SHARE_PREV: the term reads the grid slot of the
// term before it.
constexpr int FORM_XONLY = 0x10;
constexpr int FORM_SHARE_PREV = 0x20;
constexpr int base_form( int f ) { return f & 0xf; }

// Slots and jet levels of a gather pass (SHARED: one slot read by every
// term; otherwise one slot per term, except that a FORM_SHARE_PREV term
// reads its predecessor's): values for value-type outputs, gradients for
// gradient-type outputs and value-type position cotangents, Hessians for
// gradient-type position cotangents.
template <int... FORMS>
struct GatherNeeds
{
    static constexpr int N = sizeof...( FORMS );
    static constexpr int raw[N] = { FORMS... };
    static constexpr int f[N] = { base_form( FORMS )... };
    static constexpr bool xonly( int t ) { return ( raw[t] & FORM_XONLY ) != 0; }
    static constexpr int slot_of( int t, bool shared )
    {
        if ( shared )
            return 0;
        int s = -1;
        for ( int i = 0; i <= t; ++i )
            if ( !( raw[i] & FORM_SHARE_PREV ) )
                ++s;
        return s;
    }
    static constexpr int nslot( bool shared ) { return slot_of( N - 1, shared ) + 1; }
    static constexpr int first_term( int slot, bool shared )
    {
        for ( int t = 0; t < N; ++t )
            if ( slot_of( t, shared ) == slot )
                return t;
        return 0;
    }
    static constexpr bool uses( int t, int slot, bool shared )
    {
        return slot_of( t, shared ) == slot;
    }
    // The term writes its particle output / adds its position cotangent.
    static constexpr bool out( int t, int mode ) { return mode_out( mode ) && !xonly( t ); }
    static constexpr bool xbar( int mode ) { return mode_xbar( mode ); }
    static constexpr bool V( int slot, bool shared, int mode )
    {
        for ( int t = 0; t < N; ++t )
            if ( uses( t, slot, shared ) && form_value_type( f[t] ) && out( t, mode ) )
                return true;
        return false;
    }
    static constexpr bool G( int slot, bool shared, int mode )
    {
        for ( int t = 0; t < N; ++t )
            if ( uses( t, slot, shared ) &&
                 ( ( !form_value_type( f[t] ) && out( t, mode ) ) ||
                   ( form_value_type( f[t] ) && xbar( mode ) ) ) )
                return true;
        return false;
    }
    static constexpr bool H( int slot, bool shared, int mode )
    {
        for ( int t = 0; t < N; ++t )
            if ( uses( t, slot, shared ) && !form_value_type( f[t] ) && xbar( mode ) )
                return true;
        return false;
    }
    // Grid dofs of a slot: the most any of its terms reads.
    static constexpr int NG( int slot, bool shared )
    {
        int ng = 1;
        for ( int t = 0; t < N; ++t )
            if ( uses( t, slot, shared ) && !( f[t] == FORM_SV || f[t] == FORM_VD ) )
                ng = 3;
        return ng;
    }
    // Passes of more than three terms read each term's particle inputs at
    // the start of its slot (their latency hidden by that slot's jets)
    // instead of one particle ahead, which would hold every term's inputs
    // for two particles in registers.
    static constexpr bool late() { return N > 3; }
    // Position-only terms need a pass with the position cotangent.
    static constexpr bool valid( int mode )
    {
        for ( int t = 0; t < N; ++t )
            if ( xonly( t ) && !mode_xbar( mode ) )
                return false;
        return true;
    }
};

// Global reads of one particle in a gather pass: the accumulated outputs,
// the attributes the position adjoint needs, the position cotangent.
template <class Real, int MODE, int... FORMS>
struct GatherIn
{
    static constexpr int NT = sizeof...( FORMS );
    Real old[NT][9];
    Real P[NT][9];
    double xb[3];
};

template <class Real, int MODE, bool CONTIG, int... FORMS>
__device__ __forceinline__ void gather_load( const GatherArgs& a, int64_t p,
                                             GatherIn<Real, MODE, FORMS...>& in )
{
    using GN = GatherNeeds<FORMS...>;
    constexpr int NT = GN::N;
    if constexpr ( !GN::late() )
        static_for<NT>( [&]( auto I ) {
            constexpr int t = I;
            constexpr int NP = FormTraits<GN::f[t]>::NP;
            if constexpr ( GN::out( t, MODE ) )
                a.t[t].pout.template load<Real, NP, CONTIG>( p, in.old[t] );
            if constexpr ( mode_xbar( MODE ) )
                a.t[t].px.template load<Real, NP, CONTIG>( p, in.P[t] );
        } );
    if constexpr ( mode_xbar( MODE ) && !GN::late() )
        a.xbar.load<double, 3, CONTIG>( p, in.xb );
}

// One particle of a gather pass, from its global reads `in`: for each grid
// slot, its jets, then for each term reading that slot
//   pout = old + mult * dB/dP      (MODE has outputs, the term is not
//                                   position-only)
//   xbar = old + mult * dB/dx      (MODE has the position cotangent).
// Slots are finished one after the other, so only one slot's jets are live.
// Load(slot_constant, i, j, k, f) fills the slot's NG values of stencil node
// (i, j, k).
template <int ORDER, class Real, int MODE, bool CONTIG, bool SHARED, int... FORMS, class Load>
__device__ __forceinline__ void gather_compute(
    const GatherArgs& a, int64_t p, const Spline3D<ORDER, Real, mode_xbar( MODE )>& s,
    Load&& load, const GatherIn<Real, MODE, FORMS...>& in )
{
    using GN = GatherNeeds<FORMS...>;
    static_assert( GN::valid( MODE ), "position-only terms need the position cotangent" );
    constexpr int NT = GN::N;
    constexpr int NSLOT = GN::nslot( SHARED );
    constexpr bool XBAR = mode_xbar( MODE );
    double xb[3] = { 0.0, 0.0, 0.0 };
    if constexpr ( XBAR && GN::late() )
        a.xbar.load<double, 3, CONTIG>( p, xb );
    else if constexpr ( XBAR )
    {
#pragma unroll
        for ( int e = 0; e < 3; ++e )
            xb[e] = in.xb[e];
    }
    static_for<NSLOT>( [&]( auto S ) {
        constexpr int sl = S;
        constexpr int NG = GN::NG( sl, SHARED );
        // Late inputs of this slot's terms (issued before the jets).
        Real lold[NT][9];
        Real lP[NT][9];
        if constexpr ( GN::late() )
            static_for<NT>( [&]( auto I ) {
                constexpr int t = I;
                if constexpr ( GN::uses( t, sl, SHARED ) )
                {
                    constexpr int NP = FormTraits<GN::f[t]>::NP;
                    if constexpr ( GN::out( t, MODE ) )
                        a.t[t].pout.template load<Real, NP, CONTIG>( p, lold[t] );
                    if constexpr ( XBAR )
                        a.t[t].px.template load<Real, NP, CONTIG>( p, lP[t] );
                }
            } );
        Jet<Real> J[3];
        eval_jets<ORDER, NG, GN::V( sl, SHARED, MODE ), GN::G( sl, SHARED, MODE ),
                  GN::H( sl, SHARED, MODE )>(
            s, [&]( int i, int j, int k, Real* fv ) { load( S, i, j, k, fv ); }, J );
        static_for<NT>( [&]( auto I ) {
            constexpr int t = I;
            if constexpr ( GN::uses( t, sl, SHARED ) )
            {
                constexpr int F = GN::f[t];
                constexpr int NP = FormTraits<F>::NP;
                const Real m = (Real)a.t[t].mult;
                const Real* old_t = GN::late() ? lold[t] : in.old[t];
                const Real* P_t = GN::late() ? lP[t] : in.P[t];
                if constexpr ( GN::out( t, MODE ) )
                {
                    Real out[NP];
                    jet_dP<F>( J, out );
#pragma unroll
                    for ( int c = 0; c < NP; ++c )
                        out[c] = old_t[c] + m * out[c];
                    a.t[t].pout.template store<Real, NP, CONTIG>( p, out );
                }
                if constexpr ( XBAR )
                {
                    Real xt[3] = { Real( 0 ), Real( 0 ), Real( 0 ) };
                    jet_dx<F>( P_t, J, xt );
#pragma unroll
                    for ( int e = 0; e < 3; ++e )
                        xb[e] += (double)( m * xt[e] );
                }
            }
        } );
        // Finish this slot before the next slot's loads start.
        __builtin_amdgcn_sched_barrier( 0 );
    } );
    if constexpr ( XBAR )
        a.xbar.store<double, 3, CONTIG>( p, xb );
}

// One particle of a gather pass: its global reads (issued before the stencil
// work so their latency overlaps the jets), then the jets and the updates.
template <int ORDER, class Real, int MODE, bool SHARED, int... FORMS, class Load>
__device__ __forceinline__ void gather_particle(
    const GatherArgs& a, int64_t p, const Spline3D<ORDER, Real, mode_xbar( MODE )>& s,
    Load&& load )
{
    GatherIn<Real, MODE, FORMS...> in;
    gather_load<Real, MODE, false, FORMS...>( a, p, in );
    gather_compute<ORDER, Real, MODE, false, SHARED, FORMS...>( a, p, s, load, in );
}

// Loader of stencil nodes straight from the grid arrays.
template <class Real, bool SHARED, int... FORMS>
struct GlobalGatherLoad
{
    const GatherArgs& a;
    int base[3];
    template <class IC>
    __device__ __forceinline__ void operator()( IC, int i, int j, int k,
                                                Real* f ) const
    {
        using GN = GatherNeeds<FORMS...>;
        constexpr int sl = IC::value;
        constexpr int NG = GN::NG( sl, SHARED );
        const GFieldDev& g = a.t[GN::first_term( sl, SHARED )].grid;
        const Real* q = gptr<const Real>( g, ( base[0] + i ) * g.s[0] + ( base[1] + j ) * g.s[1] +
                        ( base[2] + k ) * g.s[2] );
#pragma unroll
        for ( int d = 0; d < NG; ++d )
            f[d] = q[d * g.s[3]];
    }
};

template <class Real>
__device__ __forceinline__ void atomic_add( Real* p, Real v )
{
    unsafeAtomicAdd( p, v );
}

template <bool CONTIG = false>
__device__ __forceinline__ void load_position( const PFieldDev& f, int64_t p,
                                               double px[3] )
{
    f.load<double, 3, CONTIG>( p, px );
}

} // namespace cabana_amd
