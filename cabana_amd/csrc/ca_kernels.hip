// Interpolation kernels for one spline order (CA_ORDER is set on the compile
// line; the file is compiled once per order so the three builds run in
// parallel).
//
//   gather_kernel        G2P forward and P2G adjoint: one thread per particle,
//                        K^3 stencil loads, register accumulation.
//   scatter_atomic_kernel P2G forward and G2P adjoint for particles in any
//                        order: one thread per particle, global atomics.
//   scatter_tiled_kernel P2G forward and G2P adjoint for cell-sorted
//                        particles: one workgroup per brick of cells,
//                        accumulation in an LDS node window, one flush per
//                        window node.
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

template <class Real>
__device__ __forceinline__ const Real* grid_at( const GFieldDev& g, int i,
                                                int j, int k )
{
    return reinterpret_cast<const Real*>( g.data ) + i * g.s[0] + j * g.s[1] +
           k * g.s[2];
}

// ---------------------------------------------------------------------------
// Gather: pout += mult * dB/dP, xbar += mult * dB/dx.
// ---------------------------------------------------------------------------

template <class Real, bool XBAR, bool SHARED, int... FORMS>
__global__ __launch_bounds__( BLOCK ) void gather_kernel( GatherArgs a )
{
    using FL = FormList<FORMS...>;
    constexpr int NT = FL::N;
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
    Real acc[NT][9];
    Real P[NT][9];
    Real xt[NT][3];
    const Real* gb[NT];
    static_for<NT>( [&]( auto I ) {
        constexpr int t = I;
        constexpr int NP = FormTraits<FL::v[t]>::NP;
#pragma unroll
        for ( int c = 0; c < NP; ++c )
            acc[t][c] = Real( 0 );
        xt[t][0] = xt[t][1] = xt[t][2] = Real( 0 );
        if constexpr ( XBAR )
        {
            const int64_t b = a.t[t].px.base( p );
#pragma unroll
            for ( int c = 0; c < NP; ++c )
                P[t][c] = *a.t[t].px.template ptr<const Real>( b, c );
        }
        gb[t] = grid_at<Real>( a.t[t].grid, s.d[0].base, s.d[1].base,
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
                Real Gs[3];
                if constexpr ( SHARED )
                {
                    const GFieldDev& g = a.t[0].grid;
                    const Real* gn =
                        gb[0] + i * g.s[0] + j * g.s[1] + k * g.s[2];
#pragma unroll
                    for ( int d = 0; d < 3; ++d )
                        Gs[d] = gn[d * g.s[3]];
                }
                static_for<NT>( [&]( auto I ) {
                    constexpr int t = I;
                    constexpr int FORM = FL::v[t];
                    constexpr int NG = FormTraits<FORM>::NG;
                    Real G[NG];
                    if constexpr ( SHARED )
                    {
#pragma unroll
                        for ( int d = 0; d < NG; ++d )
                            G[d] = Gs[d];
                    }
                    else
                    {
                        const GFieldDev& g = a.t[t].grid;
                        const Real* gn =
                            gb[t] + i * g.s[0] + j * g.s[1] + k * g.s[2];
#pragma unroll
                        for ( int d = 0; d < NG; ++d )
                            G[d] = gn[d * g.s[3]];
                    }
                    form_gather<FORM>( G, n, acc[t] );
                    if constexpr ( XBAR )
                        form_dx<FORM>( P[t], G, n, xt[t] );
                } );
            }
    double xb[3] = { 0.0, 0.0, 0.0 };
    static_for<NT>( [&]( auto I ) {
        constexpr int t = I;
        constexpr int NP = FormTraits<FL::v[t]>::NP;
        const Real m = (Real)a.t[t].mult;
        const PFieldDev& o = a.t[t].pout;
        if ( o.data )
        {
            const int64_t b = o.base( p );
#pragma unroll
            for ( int c = 0; c < NP; ++c )
                *o.template ptr<Real>( b, c ) += m * acc[t][c];
        }
        if constexpr ( XBAR )
        {
#pragma unroll
            for ( int e = 0; e < 3; ++e )
                xb[e] += (double)( m * xt[t][e] );
        }
    } );
    if constexpr ( XBAR )
    {
        const int64_t b = a.xbar.base( p );
#pragma unroll
        for ( int e = 0; e < 3; ++e )
            *a.xbar.ptr<double>( b, e ) += xb[e];
    }
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
        const int64_t b = a.t[t].pin.base( p );
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
        const int64_t b = a.xbar.base( p );
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
        Real* gn = reinterpret_cast<Real*>( g.data ) +
                   ( base[0] + i ) * g.s[0] + ( base[1] + j ) * g.s[1] +
                   ( base[2] + k ) * g.s[2];
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
// Tiled scatter over cell-sorted particles.
// ---------------------------------------------------------------------------

template <int... FORMS>
struct SlotLayout
{
    using FL = FormList<FORMS...>;
    // dofs per LDS slot; with SHARED only slot 0 is used (3 dofs).
    static constexpr int ng( int t ) { return form_ng( FL::v[t] ); }
    static constexpr int form_ng( int f )
    {
        return ( f == FORM_SV || f == FORM_VD ) ? 1 : 3;
    }
};

template <class Real, int... FORMS>
struct LdsSink
{
    Real* slot[sizeof...( FORMS )];
    int rel[3];
    int win1, win2;
    template <class IC>
    __device__ __forceinline__ void operator()( IC I, int i, int j, int k,
                                                const Real* G ) const
    {
        constexpr int t = IC::value;
        constexpr int NG = SlotLayout<FORMS...>::ng( t );
        const int node =
            ( ( rel[0] + i ) * win1 + ( rel[1] + j ) ) * win2 + rel[2] + k;
        Real* q = slot[t] + node * NG;
#pragma unroll
        for ( int d = 0; d < NG; ++d )
            atomicAdd( q + d, G[d] );
    }
};

template <class Real, bool XBAR, bool SHARED, int... FORMS>
__global__ __launch_bounds__( BLOCK ) void scatter_tiled_kernel( ScatterArgs a,
                                                                 TileArgs ta )
{
    using FL = FormList<FORMS...>;
    constexpr int NT = FL::N;
    constexpr int NSLOT = SHARED ? 1 : NT;
    extern __shared__ __align__( 16 ) unsigned char smem[];

    // XCD-aware mapping: consecutive blocks land on different XCDs, so
    // deal each XCD a contiguous range of tiles (neighbouring windows then
    // flush into the same L2).
    const int b = blockIdx.x;
    const int tile = ( b & 7 ) * ta.tiles_per_xcd + ( b >> 3 );
    const int ntiles = ta.ntile[0] * ta.ntile[1] * ta.ntile[2];
    if ( tile >= ntiles )
        return;
    const int ti = tile / ( ta.ntile[1] * ta.ntile[2] );
    const int tj = ( tile / ta.ntile[2] ) % ta.ntile[1];
    const int tk = tile % ta.ntile[2];
    const int c0[3] = { ti * ta.tile[0], tj * ta.tile[1], tk * ta.tile[2] };
    int c1[3];
#pragma unroll
    for ( int d = 0; d < 3; ++d )
        c1[d] = min( c0[d] + ta.tile[d], a.mesh.num_cell[d] );
    const int shift = ( ORDER == 1 ) ? 0 : -1;
    int wlo[3];
#pragma unroll
    for ( int d = 0; d < 3; ++d )
        wlo[d] = c0[d] + a.mesh.halo + shift;
    const int W0 = ta.win[0], W1 = ta.win[1], W2 = ta.win[2];
    const int wnodes = W0 * W1 * W2;

    // LDS: node windows, then the run table.
    Real* slots[NT];
    Real* lds = reinterpret_cast<Real*>( smem );
    int slot_total = 0;
    static_for<NT>( [&]( auto I ) {
        constexpr int t = I;
        if ( t < NSLOT )
        {
            slots[t] = lds + slot_total;
            slot_total += wnodes * ( SHARED ? 3 : SlotLayout<FORMS...>::ng( t ) );
        }
        else
        {
            slots[t] = lds;
        }
    } );
    int64_t* run_start =
        reinterpret_cast<int64_t*>( smem + ( ( slot_total * sizeof( Real ) + 15 ) & ~(size_t)15 ) );
    int* run_len = reinterpret_cast<int*>( run_start + ta.max_runs );
    int* max_len = run_len + ta.max_runs;

    for ( int e = threadIdx.x; e < slot_total; e += BLOCK )
        lds[e] = Real( 0 );
    const int ni = c1[0] - c0[0], nj = c1[1] - c0[1];
    const int nruns = ni * nj;
    if ( threadIdx.x == 0 )
        *max_len = 0;
    __syncthreads();
    for ( int r = threadIdx.x; r < nruns; r += BLOCK )
    {
        const int ci = c0[0] + r / nj, cj = c0[1] + r % nj;
        const int64_t cfirst =
            ( (int64_t)ci * a.mesh.num_cell[1] + cj ) * a.mesh.num_cell[2] + c0[2];
        const int64_t clast = cfirst + ( c1[2] - c0[2] ) - 1;
        const int64_t st = ta.cell_offsets[cfirst];
        const int64_t en = ta.cell_offsets[clast] + ta.cell_counts[clast];
        run_start[r] = st;
        run_len[r] = (int)( en - st );
        atomicMax( max_len, (int)( en - st ) );
    }
    __syncthreads();

    const int64_t total = (int64_t)nruns * ( *max_len );
    for ( int64_t v = threadIdx.x; v < total; v += BLOCK )
    {
        const int r = (int)( v % nruns );
        const int rank = (int)( v / nruns );
        if ( rank >= run_len[r] )
            continue;
        const int64_t p = ta.bin_begin + run_start[r] + rank;
        double px[3];
        load_position( a.points, p, px );
        Spline3D<ORDER, Real, XBAR> s;
        if ( !eval_spline( a.mesh, px, s ) )
        {
            atomicAdd( a.skipped, 1ull );
            continue;
        }
        int rel[3];
        bool fits = true;
#pragma unroll
        for ( int d = 0; d < 3; ++d )
        {
            rel[d] = s.d[d].base - wlo[d];
            fits = fits && rel[d] >= 0 && rel[d] + K <= ta.win[d];
        }
        if ( fits )
        {
            LdsSink<Real, FORMS...> sink;
            static_for<NT>( [&]( auto I ) { sink.slot[I] = slots[I]; } );
            sink.rel[0] = rel[0];
            sink.rel[1] = rel[1];
            sink.rel[2] = rel[2];
            sink.win1 = W1;
            sink.win2 = W2;
            scatter_particle<Real, XBAR, SHARED, FORMS...>( a, p, s, sink );
        }
        else
        {
            // Stencil outside this tile's window (particle not sorted into
            // this brick): add straight into global memory.
            GlobalSink<Real, FORMS...> sink{
                a, { s.d[0].base, s.d[1].base, s.d[2].base } };
            scatter_particle<Real, XBAR, SHARED, FORMS...>( a, p, s, sink );
        }
    }
    __syncthreads();

    // Flush: one global atomic per non-zero window value, walking the window
    // with the grid's unit-stride dimension fastest.
    const int pa = ta.flush_perm[0], pb = ta.flush_perm[1], pc = ta.flush_perm[2];
    const int Wd[3] = { W0, W1, W2 };
    static_for<NT>( [&]( auto I ) {
        constexpr int t = I;
        if ( t >= NSLOT )
            return;
        constexpr int NG = SHARED ? 3 : SlotLayout<FORMS...>::ng( t );
        const GFieldDev& g = a.t[t].grid;
        if ( !g.data )
            return;
        const Real* sl = slots[t];
        const int n_a = Wd[pa], n_b = Wd[pb], n_c = Wd[pc];
        const int nflush = wnodes * NG;
        for ( int f = threadIdx.x; f < nflush; f += BLOCK )
        {
            const int d = f % NG;
            const int node = f / NG;
            int q[3];
            q[pc] = node % n_c;
            q[pb] = ( node / n_c ) % n_b;
            q[pa] = node / ( n_c * n_b );
            const Real v = sl[( ( q[0] * W1 + q[1] ) * W2 + q[2] ) * NG + d];
            if ( v == Real( 0 ) )
                continue;
            const int gi = wlo[0] + q[0], gj = wlo[1] + q[1], gk = wlo[2] + q[2];
            if ( gi < 0 || gj < 0 || gk < 0 || gi >= a.mesh.num_node[0] ||
                 gj >= a.mesh.num_node[1] || gk >= a.mesh.num_node[2] )
                continue;
            Real* gp = reinterpret_cast<Real*>( g.data ) + gi * g.s[0] +
                       gj * g.s[1] + gk * g.s[2] + d * g.s[3];
            atomic_add( gp, v );
        }
    } );
}

// ---------------------------------------------------------------------------
// Launch helpers
// ---------------------------------------------------------------------------

inline dim3 blocks_for( int64_t n )
{
    return dim3( (unsigned)( ( n + BLOCK - 1 ) / BLOCK ) );
}

template <class Real, bool XBAR>
hipError_t gather_single( const GatherArgs& a, hipStream_t s )
{
    const dim3 g = blocks_for( a.end - a.begin );
    switch ( a.t[0].form )
    {
    case FORM_SV:
        hipLaunchKernelGGL( ( gather_kernel<Real, XBAR, false, FORM_SV> ), g,
                            dim3( BLOCK ), 0, s, a );
        break;
    case FORM_VV:
        hipLaunchKernelGGL( ( gather_kernel<Real, XBAR, false, FORM_VV> ), g,
                            dim3( BLOCK ), 0, s, a );
        break;
    case FORM_SG:
        hipLaunchKernelGGL( ( gather_kernel<Real, XBAR, false, FORM_SG> ), g,
                            dim3( BLOCK ), 0, s, a );
        break;
    case FORM_VD:
        hipLaunchKernelGGL( ( gather_kernel<Real, XBAR, false, FORM_VD> ), g,
                            dim3( BLOCK ), 0, s, a );
        break;
    case FORM_TD:
        hipLaunchKernelGGL( ( gather_kernel<Real, XBAR, false, FORM_TD> ), g,
                            dim3( BLOCK ), 0, s, a );
        break;
    case FORM_TDT:
        hipLaunchKernelGGL( ( gather_kernel<Real, XBAR, false, FORM_TDT> ), g,
                            dim3( BLOCK ), 0, s, a );
        break;
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
    const char* name =
        a.has_xbar ? ( plan == PLAN_MPM_P2G ? "p2g_adjoint_mpm" : "p2g_adjoint" )
                   : ( plan == PLAN_MPM_G2P ? "g2p_mpm" : "g2p" );
    L.before( name, s );
    if ( plan == PLAN_SINGLE )
    {
        e = a.has_xbar ? gather_single<Real, true>( a, s )
                       : gather_single<Real, false>( a, s );
    }
    else if ( plan == PLAN_MPM_G2P && !a.has_xbar )
    {
        hipLaunchKernelGGL(
            ( gather_kernel<Real, false, true, FORM_VV, FORM_TDT> ), g,
            dim3( BLOCK ), 0, s, a );
        e = hipGetLastError();
    }
    else if ( plan == PLAN_MPM_P2G && a.has_xbar )
    {
        hipLaunchKernelGGL(
            ( gather_kernel<Real, true, false, FORM_SV, FORM_VV, FORM_TD> ), g,
            dim3( BLOCK ), 0, s, a );
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
            : ( plan == PLAN_MPM_P2G ? "p2g_mpm_atomic" : "p2g_atomic" );
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
    else if ( plan == PLAN_MPM_G2P && a.has_xbar )
    {
        hipLaunchKernelGGL(
            ( scatter_atomic_kernel<Real, true, true, FORM_VV, FORM_TDT> ), g,
            dim3( BLOCK ), 0, s, a );
        e = hipGetLastError();
    }
    else
    {
        e = hipErrorInvalidValue;
    }
    L.after( name, s );
    return e;
}

template <class KernelPtr>
hipError_t launch_tiled( KernelPtr k, const ScatterArgs& a, const TileArgs& t,
                         size_t lds, hipStream_t s )
{
    if ( lds > 64 * 1024 )
    {
        hipError_t e = hipFuncSetAttribute(
            reinterpret_cast<const void*>( k ),
            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds );
        if ( e != hipSuccess )
            return e;
    }
    const dim3 g( (unsigned)( 8 * t.tiles_per_xcd ) );
    hipLaunchKernelGGL( k, g, dim3( BLOCK ), lds, s, a, t );
    return hipGetLastError();
}

template <class Real>
size_t tiled_lds_bytes( int plan, const ScatterArgs& a, const TileArgs& t )
{
    const size_t wn = (size_t)t.win[0] * t.win[1] * t.win[2];
    size_t dofs = 0;
    if ( plan == PLAN_MPM_G2P )
        dofs = 3;
    else
        for ( int i = 0; i < a.nterms; ++i )
            dofs += ( a.t[i].form == FORM_SV || a.t[i].form == FORM_VD ) ? 1 : 3;
    size_t bytes = ( ( wn * dofs * sizeof( Real ) ) + 15 ) & ~(size_t)15;
    bytes += (size_t)t.max_runs * ( sizeof( int64_t ) + sizeof( int ) ) + 16;
    return bytes;
}

template <class Real>
hipError_t scatter_tiled_typed( Launcher& L, int plan, const ScatterArgs& a,
                                const TileArgs& t, hipStream_t s )
{
    const size_t lds = tiled_lds_bytes<Real>( plan, a, t );
    hipError_t e = hipSuccess;
    const char* name =
        a.has_xbar
            ? ( plan == PLAN_MPM_G2P ? "g2p_adjoint_mpm_tiled" : "g2p_adjoint_tiled" )
            : ( plan == PLAN_MPM_P2G ? "p2g_mpm_tiled" : "p2g_tiled" );
    L.before( name, s );
    if ( plan == PLAN_SINGLE )
    {
        switch ( a.t[0].form )
        {
#define CA_CASE( F )                                                           \
    case F:                                                                    \
        e = a.has_xbar                                                         \
                ? launch_tiled( scatter_tiled_kernel<Real, true, false, F>, a, \
                                t, lds, s )                                    \
                : launch_tiled( scatter_tiled_kernel<Real, false, false, F>,   \
                                a, t, lds, s );                                \
        break;
            CA_CASE( FORM_SV )
            CA_CASE( FORM_VV )
            CA_CASE( FORM_SG )
            CA_CASE( FORM_VD )
            CA_CASE( FORM_TD )
            CA_CASE( FORM_TDT )
#undef CA_CASE
        default:
            e = hipErrorInvalidValue;
        }
    }
    else if ( plan == PLAN_MPM_P2G && !a.has_xbar )
    {
        e = launch_tiled(
            scatter_tiled_kernel<Real, false, false, FORM_SV, FORM_VV, FORM_TD>,
            a, t, lds, s );
    }
    else if ( plan == PLAN_MPM_G2P && a.has_xbar )
    {
        e = launch_tiled(
            scatter_tiled_kernel<Real, true, true, FORM_VV, FORM_TDT>, a, t,
            lds, s );
    }
    else
    {
        e = hipErrorInvalidValue;
    }
    L.after( name, s );
    return e;
}

} // namespace

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

hipError_t CA_CAT( launch_scatter_tiled_, CA_SUFFIX )( Launcher& L, bool fp64,
                                                      int plan,
                                                      const ScatterArgs& a,
                                                      const TileArgs& t,
                                                      hipStream_t s )
{
    return fp64 ? scatter_tiled_typed<double>( L, plan, a, t, s )
                : scatter_tiled_typed<float>( L, plan, a, t, s );
}

} // namespace cabana_amd
