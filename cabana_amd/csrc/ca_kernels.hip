// Interpolation kernels for one spline order (CA_ORDER is set on the compile
// line; the file is compiled once per order so the three builds run in
// parallel).
//
//   gather_kernel        G2P forward and P2G adjoint: one thread per particle,
//                        K^3 stencil loads, register accumulation.
//   scatter_atomic_kernel P2G forward and G2P adjoint for particles in any
//                        order: one thread per particle, global atomics.
// The cell-sorted (tiled) variants live in ca_tiled.hip.
#include "ca_device.hpp"

#include <utility>

#ifndef CA_ORDER
#error "CA_ORDER must be defined"
#endif

#define CA_CAT2( a, b ) a##b
#define CA_CAT( a, b ) CA_CAT2( a, b )
#define CA_SUFFIX CA_CAT( o, CA_ORDER )

namespace cabana_amd
{
namespace
{

constexpr int ORDER = CA_ORDER;
constexpr int K = ORDER + 1;
constexpr int BLOCK = 256;

template <int... FORMS>
struct FormList
{
    static constexpr int N = sizeof...( FORMS );
    static constexpr int v[N] = { FORMS... };
};

template <class Real>
__device__ __forceinline__ const Real* grid_at( const GFieldDev& g, int i,
                                                int j, int k )
{
    return gptr<const Real>( g, i * g.s[0] + j * g.s[1] +
           k * g.s[2] );
}

// ---------------------------------------------------------------------------
// Gather: pout += mult * dB/dP, xbar += mult * dB/dx.
// ---------------------------------------------------------------------------

template <class Real, int MODE, bool SHARED, int... FORMS>
__global__ __launch_bounds__( BLOCK ) void gather_kernel(
    GatherArgs a )
{
    const int64_t p = a.begin + (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if ( p >= a.end )
        return;
    double px[3];
    load_position( a.points, p, px );
    Spline3D<ORDER, Real, mode_xbar( MODE )> s;
    if ( !eval_spline( a.mesh, px, s ) )
    {
        atomicAdd( a.skipped, 1ull );
        return;
    }
    GlobalGatherLoad<Real, SHARED, FORMS...> ld{
        a, { s.d[0].base, s.d[1].base, s.d[2].base } };
    gather_particle<ORDER, Real, MODE, SHARED, FORMS...>( a, p, s, ld );
}

// ---------------------------------------------------------------------------
// Scatter building block: contributions of one particle for every term.
// Sink(t, i, j, k, G[NG]) receives the node contribution of term t.
// ---------------------------------------------------------------------------

template <class Real, bool XBAR, bool SHARED, int... FORMS, class Sink>
__device__ __forceinline__ void scatter_particle( const ScatterArgs& a,
                                                  int64_t p,
                                                  const Spline3D<ORDER, Real, XBAR>& s,
                                                  Sink&& sink )
{
    using FL = FormList<FORMS...>;
    constexpr int NT = FL::N;
    Real P[NT][9];
    Real Pm[NT][9];
    Real xt[NT][3];
    const Real* xg[NT];
    static_for<NT>( [&]( auto I ) {
        constexpr int t = I;
        constexpr int NP = FormTraits<FL::v[t]>::NP;
        const Real m = (Real)a.t[t].mult;
        const uint32_t b = a.t[t].pin.base( p );
#pragma unroll
        for ( int c = 0; c < NP; ++c )
        {
            P[t][c] = *a.t[t].pin.template ptr<const Real>( b, c );
            Pm[t][c] = P[t][c] * m;
        }
        xt[t][0] = xt[t][1] = xt[t][2] = Real( 0 );
        if constexpr ( XBAR )
            xg[t] = grid_at<Real>( a.t[t].gx, s.d[0].base, s.d[1].base,
                                   s.d[2].base );
    } );
#pragma unroll
    for ( int i = 0; i < K; ++i )
#pragma unroll
        for ( int j = 0; j < K; ++j )
#pragma unroll
            for ( int k = 0; k < K; ++k )
            {
                const auto n =
                    node_weights<ORDER, Real, XBAR, XBAR>( s, i, j, k );
                Real Gsum[3] = { Real( 0 ), Real( 0 ), Real( 0 ) };
                Real Gx[3];
                if constexpr ( XBAR && SHARED )
                {
                    const GFieldDev& g = a.t[0].gx;
                    const Real* gn =
                        xg[0] + i * g.s[0] + j * g.s[1] + k * g.s[2];
#pragma unroll
                    for ( int d = 0; d < 3; ++d )
                        Gx[d] = gn[d * g.s[3]];
                }
                static_for<NT>( [&]( auto I ) {
                    constexpr int t = I;
                    constexpr int FORM = FL::v[t];
                    constexpr int NG = FormTraits<FORM>::NG;
                    Real G[NG];
                    form_scatter<FORM>( Pm[t], n, G );
                    if constexpr ( SHARED )
                    {
#pragma unroll
                        for ( int d = 0; d < NG; ++d )
                            Gsum[d] += G[d];
                    }
                    else
                    {
                        sink( I, i, j, k, G );
                    }
                    if constexpr ( XBAR )
                    {
                        Real Gr[NG];
                        if constexpr ( SHARED )
                        {
#pragma unroll
                            for ( int d = 0; d < NG; ++d )
                                Gr[d] = Gx[d];
                        }
                        else
                        {
                            const GFieldDev& g = a.t[t].gx;
                            const Real* gn =
                                xg[t] + i * g.s[0] + j * g.s[1] + k * g.s[2];
#pragma unroll
                            for ( int d = 0; d < NG; ++d )
                                Gr[d] = gn[d * g.s[3]];
                        }
                        form_dx<FORM>( P[t], Gr, n, xt[t] );
                    }
                } );
                if constexpr ( SHARED )
                    sink( std::integral_constant<int, 0>{}, i, j, k, Gsum );
            }
    if constexpr ( XBAR )
    {
        double xb[3] = { 0.0, 0.0, 0.0 };
        static_for<NT>( [&]( auto I ) {
            constexpr int t = I;
            const Real m = (Real)a.t[t].mult;
#pragma unroll
            for ( int e = 0; e < 3; ++e )
                xb[e] += (double)( m * xt[t][e] );
        } );
        const uint32_t b = a.xbar.base( p );
#pragma unroll
        for ( int e = 0; e < 3; ++e )
            *a.xbar.ptr<double>( b, e ) += xb[e];
    }
}

// Global-atomic sink: adds one node contribution into term t's grid.
template <class Real, int... FORMS>
struct GlobalSink
{
    const ScatterArgs& a;
    int base[3];
    template <class IC>
    __device__ __forceinline__ void operator()( IC I, int i, int j, int k,
                                                const Real* G ) const
    {
        using FL = FormList<FORMS...>;
        constexpr int NG = FormTraits<FL::v[IC::value]>::NG;
        const GFieldDev& g = a.t[IC::value].grid;
        if ( !g.data )
            return;
        Real* gn = gptr<Real>( g, ( base[0] + i ) * g.s[0] + ( base[1] + j ) * g.s[1] +
                   ( base[2] + k ) * g.s[2] );
#pragma unroll
        for ( int d = 0; d < NG; ++d )
            atomic_add( gn + d * g.s[3], G[d] );
    }
};

template <class Real, bool XBAR, bool SHARED, int... FORMS>
__global__ __launch_bounds__( BLOCK ) void scatter_atomic_kernel( ScatterArgs a )
{
    const int64_t p = a.begin + (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if ( p >= a.end )
        return;
    double px[3];
    load_position( a.points, p, px );
    Spline3D<ORDER, Real, XBAR> s;
    if ( !eval_spline( a.mesh, px, s ) )
    {
        atomicAdd( a.skipped, 1ull );
        return;
    }
    GlobalSink<Real, FORMS...> sink{ a,
                                     { s.d[0].base, s.d[1].base, s.d[2].base } };
    scatter_particle<Real, XBAR, SHARED, FORMS...>( a, p, s, sink );
}

// ---------------------------------------------------------------------------
// Any set of terms in one pass (the reference benchmark's fused_p2g,
// fused_g2p and g2p2g, Cabana_Grid_InterpolationPerformance.cpp:25-174)
// ---------------------------------------------------------------------------

// One G2P term of one particle: pout += mult * dB/dP over the stencil.
template <class Real, int FORM>
__device__ __forceinline__ void gather_term( const MultiTermDev& term, int64_t p,
                                             const Spline3D<ORDER, Real, false>& s )
{
    constexpr int NP = FormTraits<FORM>::NP;
    constexpr int NG = FormTraits<FORM>::NG;
    const GFieldDev& g = term.grid;
    const Real* g0 = grid_at<Real>( g, s.d[0].base, s.d[1].base, s.d[2].base );
    Real acc[NP];
#pragma unroll
    for ( int c = 0; c < NP; ++c )
        acc[c] = Real( 0 );
#pragma unroll
    for ( int i = 0; i < K; ++i )
#pragma unroll
        for ( int j = 0; j < K; ++j )
#pragma unroll
            for ( int k = 0; k < K; ++k )
            {
                const auto n = node_weights<ORDER, Real, false, false>( s, i, j, k );
                const Real* gn = g0 + i * g.s[0] + j * g.s[1] + k * g.s[2];
                Real G[NG];
#pragma unroll
                for ( int d = 0; d < NG; ++d )
                    G[d] = gn[d * g.s[3]];
                form_gather<FORM>( G, n, acc );
            }
    const Real m = (Real)term.mult;
    const uint32_t b = term.f.base( p );
#pragma unroll
    for ( int c = 0; c < NP; ++c )
    {
        Real* o = term.f.template ptr<Real>( b, c );
        *o += m * acc[c];
    }
}

// One P2G term of one particle: grid += dB/dG of mult * P over the stencil.
template <class Real, int FORM>
__device__ __forceinline__ void scatter_term( const MultiTermDev& term, int64_t p,
                                              const Spline3D<ORDER, Real, false>& s )
{
    constexpr int NP = FormTraits<FORM>::NP;
    constexpr int NG = FormTraits<FORM>::NG;
    const Real m = (Real)term.mult;
    const uint32_t b = term.f.base( p );
    Real Pm[NP];
#pragma unroll
    for ( int c = 0; c < NP; ++c )
        Pm[c] = *term.f.template ptr<const Real>( b, c ) * m;
    const GFieldDev& g = term.grid;
    Real* g0 = gptr<Real>( g, s.d[0].base * g.s[0] + s.d[1].base * g.s[1] +
                                  s.d[2].base * g.s[2] );
#pragma unroll
    for ( int i = 0; i < K; ++i )
#pragma unroll
        for ( int j = 0; j < K; ++j )
#pragma unroll
            for ( int k = 0; k < K; ++k )
            {
                const auto n = node_weights<ORDER, Real, false, false>( s, i, j, k );
                Real G[NG];
                form_scatter<FORM>( Pm, n, G );
                Real* gn = g0 + i * g.s[0] + j * g.s[1] + k * g.s[2];
#pragma unroll
                for ( int d = 0; d < NG; ++d )
                    atomic_add( gn + d * g.s[3], G[d] );
            }
}

template <class Real>
__global__ __launch_bounds__( BLOCK ) void multi_term_kernel( MultiTermArgs a )
{
    const int64_t p = a.begin + (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if ( p >= a.end )
        return;
    double px[3];
    load_position( a.points, p, px );
    Spline3D<ORDER, Real, false> s;
    if ( !eval_spline( a.mesh, px, s ) )
    {
        atomicAdd( a.skipped, 1ull );
        return;
    }
    // The term index is uniform across the wave, so each switch is a
    // scalar branch.
    for ( int t = 0; t < a.n_g2p; ++t )
    {
        switch ( a.g[t].form )
        {
#define CA_CASE( F )                                                           \
    case F:                                                                    \
        gather_term<Real, F>( a.g[t], p, s );                                  \
        break;
            CA_CASE( FORM_SV )
            CA_CASE( FORM_VV )
            CA_CASE( FORM_SG )
            CA_CASE( FORM_VD )
            CA_CASE( FORM_TDT )
#undef CA_CASE
        default:
            break;
        }
    }
    for ( int t = 0; t < a.n_p2g; ++t )
    {
        switch ( a.s[t].form )
        {
#define CA_CASE( F )                                                           \
    case F:                                                                    \
        scatter_term<Real, F>( a.s[t], p, s );                                 \
        break;
            CA_CASE( FORM_SV )
            CA_CASE( FORM_VV )
            CA_CASE( FORM_SG )
            CA_CASE( FORM_VD )
            CA_CASE( FORM_TD )
#undef CA_CASE
        default:
            break;
        }
    }
}

// ---------------------------------------------------------------------------
// Launch helpers
// ---------------------------------------------------------------------------

inline dim3 blocks_for( int64_t n )
{
    return dim3( (unsigned)( ( n + BLOCK - 1 ) / BLOCK ) );
}

// Gather mode of a pass: outputs and/or the position cotangent (a fused
// pass has either every term's output or none).
inline int gather_mode( const GatherArgs& a )
{
    if ( !a.has_xbar )
        return GATHER_OUT;
    return a.t[0].pout.data ? GATHER_OUT_XBAR : GATHER_XBAR;
}

template <class Real, int MODE>
hipError_t gather_single( const GatherArgs& a, hipStream_t s )
{
    const dim3 g = blocks_for( a.end - a.begin );
    switch ( a.t[0].form )
    {
#define CA_CASE( F )                                                           \
    case F:                                                                    \
        hipLaunchKernelGGL( ( gather_kernel<Real, MODE, false, F> ), g, dim3( BLOCK ), 0, \
                            s, a );                                            \
        break;
        CA_CASE( FORM_SV )
        CA_CASE( FORM_VV )
        CA_CASE( FORM_SG )
        CA_CASE( FORM_VD )
        CA_CASE( FORM_TD )
        CA_CASE( FORM_TDT )
#undef CA_CASE
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <class Real>
hipError_t gather_typed( Launcher& L, int plan, const GatherArgs& a,
                         hipStream_t s )
{
    if ( a.end <= a.begin )
        return hipSuccess;
    const dim3 g = blocks_for( a.end - a.begin );
    hipError_t e = hipSuccess;
    const int mode = gather_mode( a );
    const char* name =
        plan == PLAN_MPM_ADJOINT
            ? "adjoint_mpm"
            : a.has_xbar ? ( plan == PLAN_MPM_P2G ? "p2g_adjoint_mpm" : "p2g_adjoint" )
                         : ( plan == PLAN_MPM_G2P ? "g2p_mpm" : "g2p" );
    L.before( name, s );
    if ( plan == PLAN_SINGLE )
    {
        e = mode == GATHER_OUT        ? gather_single<Real, GATHER_OUT>( a, s )
            : mode == GATHER_OUT_XBAR ? gather_single<Real, GATHER_OUT_XBAR>( a, s )
                                      : gather_single<Real, GATHER_XBAR>( a, s );
    }
    else if ( plan == PLAN_MPM_G2P && mode == GATHER_OUT )
    {
        hipLaunchKernelGGL(
            ( gather_kernel<Real, GATHER_OUT, true, FORM_VV, FORM_TDT> ), g,
            dim3( BLOCK ), 0, s, a );
        e = hipGetLastError();
    }
    else if ( plan == PLAN_MPM_G2P && mode == GATHER_XBAR )
    {
        hipLaunchKernelGGL(
            ( gather_kernel<Real, GATHER_XBAR, true, FORM_VV, FORM_TDT> ), g,
            dim3( BLOCK ), 0, s, a );
        e = hipGetLastError();
    }
    else if ( plan == PLAN_MPM_P2G && mode == GATHER_OUT_XBAR )
    {
        hipLaunchKernelGGL(
            ( gather_kernel<Real, GATHER_OUT_XBAR, false, FORM_SV, FORM_VV, FORM_TD> ),
            g, dim3( BLOCK ), 0, s, a );
        e = hipGetLastError();
    }
    else if ( plan == PLAN_MPM_ADJOINT && mode == GATHER_OUT_XBAR )
    {
        hipLaunchKernelGGL(
            ( gather_kernel<Real, GATHER_OUT_XBAR, false, FORM_SV, FORM_VV, FORM_TD,
                            FORM_VV | FORM_XONLY, FORM_TDT | FORM_XONLY | FORM_SHARE_PREV> ),
            g, dim3( BLOCK ), 0, s, a );
        e = hipGetLastError();
    }
    else
    {
        e = hipErrorInvalidValue;
    }
    L.after( name, s );
    return e;
}

template <class Real, bool XBAR>
hipError_t scatter_atomic_single( const ScatterArgs& a, hipStream_t s )
{
    const dim3 g = blocks_for( a.end - a.begin );
    switch ( a.t[0].form )
    {
#define CA_CASE( F )                                                           \
    case F:                                                                    \
        hipLaunchKernelGGL( ( scatter_atomic_kernel<Real, XBAR, false, F> ),   \
                            g, dim3( BLOCK ), 0, s, a );                       \
        break;
        CA_CASE( FORM_SV )
        CA_CASE( FORM_VV )
        CA_CASE( FORM_SG )
        CA_CASE( FORM_VD )
        CA_CASE( FORM_TD )
        CA_CASE( FORM_TDT )
#undef CA_CASE
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <class Real>
hipError_t scatter_atomic_typed( Launcher& L, int plan, const ScatterArgs& a,
                                 hipStream_t s )
{
    if ( a.end <= a.begin )
        return hipSuccess;
    const dim3 g = blocks_for( a.end - a.begin );
    hipError_t e = hipSuccess;
    const char* name =
        a.has_xbar
            ? ( plan == PLAN_MPM_G2P ? "g2p_adjoint_mpm_atomic" : "g2p_adjoint_atomic" )
            : ( plan == PLAN_MPM_P2G         ? "p2g_mpm_atomic"
                : plan == PLAN_MASS_MOMENTUM ? "p2g_mass_momentum_atomic"
                                             : "p2g_atomic" );
    L.before( name, s );
    if ( plan == PLAN_SINGLE )
    {
        e = a.has_xbar ? scatter_atomic_single<Real, true>( a, s )
                       : scatter_atomic_single<Real, false>( a, s );
    }
    else if ( plan == PLAN_MPM_P2G && !a.has_xbar )
    {
        hipLaunchKernelGGL( ( scatter_atomic_kernel<Real, false, false, FORM_SV,
                                                    FORM_VV, FORM_TD> ),
                            g, dim3( BLOCK ), 0, s, a );
        e = hipGetLastError();
    }
    else if ( plan == PLAN_MASS_MOMENTUM && !a.has_xbar )
    {
        hipLaunchKernelGGL( ( scatter_atomic_kernel<Real, false, false, FORM_SV, FORM_VV> ),
                            g, dim3( BLOCK ), 0, s, a );
        e = hipGetLastError();
    }
    else if ( plan == PLAN_MPM_G2P )
    {
        if ( a.has_xbar )
            hipLaunchKernelGGL(
                ( scatter_atomic_kernel<Real, true, true, FORM_VV, FORM_TDT> ),
                g, dim3( BLOCK ), 0, s, a );
        else
            hipLaunchKernelGGL(
                ( scatter_atomic_kernel<Real, false, true, FORM_VV, FORM_TDT> ),
                g, dim3( BLOCK ), 0, s, a );
        e = hipGetLastError();
    }
    else
    {
        e = hipErrorInvalidValue;
    }
    L.after( name, s );
    return e;
}

// ---------------------------------------------------------------------------
// Deterministic P2G ("pull"): one thread per node of the grid
// ---------------------------------------------------------------------------

// Weight and gradient of stencil node j (0 <= j < K) of a 1-D spline, the
// node chosen at run time from compile-time unrolled candidates.
template <class Real>
__device__ __forceinline__ void pick_node( const Spline1D<ORDER, Real>& s, int j, Real& w,
                                           Real& g )
{
    w = Real( 0 );
    g = Real( 0 );
    static_for<K>( [&]( auto J ) {
        constexpr int jj = J;
        w = ( j == jj ) ? s.w[jj] : w;
        g = ( j == jj ) ? s.g[jj] : g;
    } );
}

template <class Real, int... FORMS>
__global__ __launch_bounds__( BLOCK ) void p2g_pull_kernel( PullArgs a )
{
    using FL = FormList<FORMS...>;
    constexpr int NT = FL::N;
    const MeshDev& m = a.mesh;
    const GFieldDev& g = a.term[0].grid;
    const int64_t total = (int64_t)g.ext[0] * g.ext[1] * g.ext[2];
    const int64_t e = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if ( e >= total )
        return;
    const int nd[3] = { (int)( e / ( (int64_t)g.ext[1] * g.ext[2] ) ),
                        (int)( ( e / g.ext[2] ) % g.ext[1] ), (int)( e % g.ext[2] ) };
    // A particle of owned cell c has its stencil base at c + h - ORDER / 2
    // (or one more, for the dual grid of even orders) and reaches K nodes, so
    // node n sees cells [n - h + ORDER / 2 - K, n - h + ORDER / 2].  Particles
    // outside the owned cells are binned into the boundary cell nearest to
    // them, so the range is clamped into the owned cells without becoming
    // empty; every candidate is checked against its actual stencil.
    int lo[3], hi[3];
#pragma unroll
    for ( int d = 0; d < 3; ++d )
    {
        const int N = m.num_cell[d];
        const int top = nd[d] - m.halo + ORDER / 2;
        lo[d] = min( max( top - K, 0 ), N - 1 );
        hi[d] = min( max( top, 0 ), N - 1 );
    }
    double acc[NT][3];
#pragma unroll
    for ( int t = 0; t < NT; ++t )
#pragma unroll
        for ( int c = 0; c < 3; ++c )
            acc[t][c] = 0.0;
    for ( int ci = lo[0]; ci <= hi[0]; ++ci )
        for ( int cj = lo[1]; cj <= hi[1]; ++cj )
            for ( int ck = lo[2]; ck <= hi[2]; ++ck )
            {
                const int64_t cell = ( (int64_t)ci * m.num_cell[1] + cj ) * m.num_cell[2] + ck;
                const int64_t q0 = a.offsets[cell];
                const int64_t q1 = q0 + a.counts[cell];
                // The next candidate's index and position are read while
                // this one is tested and added.
                int64_t pn = 0;
                double pxn[3] = { 0.0, 0.0, 0.0 };
                if ( q0 < q1 )
                {
                    pn = a.perm[q0];
                    load_position( a.points, pn, pxn );
                }
                for ( int64_t q = q0; q < q1; ++q )
                {
                    const int64_t p = pn;
                    double px[3] = { pxn[0], pxn[1], pxn[2] };
                    if ( q + 1 < q1 )
                    {
                        pn = a.perm[q + 1];
                        load_position( a.points, pn, pxn );
                    }
                    // Stencil test first (most candidates miss the node),
                    // weights only for the particles that reach it.
                    Real f[3];
                    int j[3];
                    bool hit = true;
#pragma unroll
                    for ( int d = 0; d < 3; ++d )
                    {
                        const double x0 = logical_coord<ORDER>( px[d], m.low[d], m.rdx[d] );
                        int base = 0;
                        const bool in = spline_frac<ORDER, Real>( x0, m.num_node[d], base, f[d] );
                        j[d] = nd[d] - base;
                        hit = hit && in && j[d] >= 0 && j[d] < K;
                    }
                    if ( !hit )
                        continue;
                    Real w[3], gr[3];
#pragma unroll
                    for ( int d = 0; d < 3; ++d )
                    {
                        Spline1D<ORDER, Real> s1;
                        spline_weights<ORDER, Real, false>( f[d], (Real)m.rdx[d], s1 );
                        pick_node( s1, j[d], w[d], gr[d] );
                    }
                    NodeW<Real, false> nw;
                    const Real wyz = w[1] * w[2];
                    nw.W = w[0] * wyz;
                    nw.dW[0] = gr[0] * wyz;
                    nw.dW[1] = w[0] * gr[1] * w[2];
                    nw.dW[2] = w[0] * w[1] * gr[2];
                    static_for<NT>( [&]( auto I ) {
                        constexpr int t = I;
                        constexpr int F = FL::v[t];
                        constexpr int NP = FormTraits<F>::NP;
                        constexpr int NG = FormTraits<F>::NG;
                        const Real mult = (Real)a.term[t].mult;
                        const uint32_t b = a.term[t].f.base( p );
                        Real Pm[NP];
#pragma unroll
                        for ( int c = 0; c < NP; ++c )
                            Pm[c] = *a.term[t].f.template ptr<const Real>( b, c ) * mult;
                        Real G[NG];
                        form_scatter<F>( Pm, nw, G );
#pragma unroll
                        for ( int c = 0; c < NG; ++c )
                            acc[t][c] += (double)G[c];
                    } );
                }
            }
    static_for<NT>( [&]( auto I ) {
        constexpr int t = I;
        constexpr int NG = FormTraits<FL::v[t]>::NG;
        const GFieldDev& gt = a.term[t].grid;
        Real* dst = gptr<Real>( gt, nd[0] * gt.s[0] + nd[1] * gt.s[1] + nd[2] * gt.s[2] );
#pragma unroll
        for ( int c = 0; c < NG; ++c )
            if ( acc[t][c] != 0.0 )
                dst[c * gt.s[3]] += (Real)acc[t][c];
    } );
}

// Particles whose stencil leaves the ghosted grid (counted once each, as
// the other paths count them).
__global__ __launch_bounds__( BLOCK ) void count_outside_kernel( MeshDev m, PFieldDev pts,
                                                                 int64_t n,
                                                                 unsigned long long* skipped )
{
    const int64_t p = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if ( p >= n )
        return;
    double px[3];
    load_position( pts, p, px );
    int b[3];
    const bool out = !stencil_base<ORDER>( m, px, b );
    const unsigned long long mask = __ballot( out );
    if ( mask && (int)( threadIdx.x & 63 ) == __ffsll( (long long)mask ) - 1 )
        atomicAdd( skipped, (unsigned long long)__popcll( mask ) );
}

template <class Real>
hipError_t pull_typed( Launcher& L, const PullArgs& a, hipStream_t s )
{
    const int64_t total =
        (int64_t)a.term[0].grid.ext[0] * a.term[0].grid.ext[1] * a.term[0].grid.ext[2];
    if ( total <= 0 )
        return hipSuccess;
    L.before( a.nterms == 3 ? "p2g_pull_mpm" : "p2g_pull", s );
    if ( a.nterms == 3 )
    {
        if ( a.term[0].form != FORM_SV || a.term[1].form != FORM_VV || a.term[2].form != FORM_TD )
            return hipErrorInvalidValue;
        hipLaunchKernelGGL( ( p2g_pull_kernel<Real, FORM_SV, FORM_VV, FORM_TD> ),
                            blocks_for( total ), dim3( BLOCK ), 0, s, a );
        L.after( "p2g_pull_mpm", s );
        return hipGetLastError();
    }
    if ( a.nterms != 1 )
        return hipErrorInvalidValue;
    switch ( a.term[0].form )
    {
#define CA_CASE( F )                                                           \
    case F:                                                                    \
        hipLaunchKernelGGL( ( p2g_pull_kernel<Real, F> ), blocks_for( total ), dim3( BLOCK ), 0, \
                            s, a );                                            \
        break;
        CA_CASE( FORM_SV )
        CA_CASE( FORM_VV )
        CA_CASE( FORM_SG )
        CA_CASE( FORM_VD )
        CA_CASE( FORM_TD )
#undef CA_CASE
    default:
        return hipErrorInvalidValue;
    }
    L.after( "p2g_pull", s );
    return hipGetLastError();
}

template <class Real>
hipError_t multi_term_typed( Launcher& L, const MultiTermArgs& a, hipStream_t s )
{
    if ( a.end <= a.begin )
        return hipSuccess;
    const char* name = a.n_g2p == 0 ? "p2g_terms" : ( a.n_p2g == 0 ? "g2p_terms" : "g2p2g" );
    L.before( name, s );
    hipLaunchKernelGGL( multi_term_kernel<Real>, blocks_for( a.end - a.begin ), dim3( BLOCK ), 0,
                        s, a );
    L.after( name, s );
    return hipGetLastError();
}

} // namespace

hipError_t CA_CAT( launch_p2g_pull_, CA_SUFFIX )( Launcher& L, bool fp64, const PullArgs& a,
                                                 hipStream_t s )
{
    return fp64 ? pull_typed<double>( L, a, s ) : pull_typed<float>( L, a, s );
}

hipError_t CA_CAT( launch_count_outside_, CA_SUFFIX )( Launcher& L, const MeshDev& m,
                                                      const PFieldDev& pts, int64_t n,
                                                      unsigned long long* skipped,
                                                      hipStream_t s )
{
    if ( n <= 0 )
        return hipSuccess;
    L.before( "count_outside", s );
    hipLaunchKernelGGL( count_outside_kernel, blocks_for( n ), dim3( BLOCK ), 0, s, m, pts, n,
                        skipped );
    L.after( "count_outside", s );
    return hipGetLastError();
}

hipError_t CA_CAT( launch_multi_term_, CA_SUFFIX )( Launcher& L, bool fp64,
                                                   const MultiTermArgs& a, hipStream_t s )
{
    return fp64 ? multi_term_typed<double>( L, a, s ) : multi_term_typed<float>( L, a, s );
}

hipError_t CA_CAT( launch_gather_, CA_SUFFIX )( Launcher& L, bool fp64,
                                               int plan, const GatherArgs& a,
                                               hipStream_t s )
{
    return fp64 ? gather_typed<double>( L, plan, a, s )
                : gather_typed<float>( L, plan, a, s );
}

hipError_t CA_CAT( launch_scatter_atomic_, CA_SUFFIX )( Launcher& L, bool fp64,
                                                       int plan,
                                                       const ScatterArgs& a,
                                                       hipStream_t s )
{
    return fp64 ? scatter_atomic_typed<double>( L, plan, a, s )
                : scatter_atomic_typed<float>( L, plan, a, s );
}

} // namespace cabana_amd
