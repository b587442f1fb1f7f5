// Cell-tiled interpolation kernels for cell-sorted particles, one spline
// order per translation unit (CA_ORDER set on the compile line).
//
// Both kernels give one workgroup a brick of T^3 mesh cells.  The brick's
// particles are the T^2 k-runs of the cell-sorted list (cell offsets of the
// binning), walked through a flattened run index so consecutive threads read
// consecutive particles.
//
//   gather_tiled_kernel   G2P forward, P2G adjoint and the position part of
//                         the G2P adjoint.  The node window of every grid the
//                         pass reads ((T + K - 1 + shift)^3 nodes) is staged
//                         in LDS once with coalesced loads; every particle
//                         then gathers its K^3 stencil from LDS.
//   scatter_cells_kernel  P2G forward and the grid part of the G2P adjoint.
//                         The brick's particles are counting-sorted in LDS by
//                         their stencil base node; one lane owns one
//                         (base cell, x-slab) pair, accumulates the slab's
//                         K^2 nodes in registers over all particles of the
//                         base cell, and flushes once into an FP64 LDS window
//                         (LDS FP64 adds are ~16x faster than FP32 ones on
//                         gfx950).  The window is flushed into the grid once.
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
// Stencil base node of a particle in cell c (ghosted node index):
// c + halo + BASE_SHIFT + {0, 1 if ORDER == 2}.
constexpr int BASE_SHIFT = ( ORDER == 1 ) ? 0 : -1;
constexpr int BASE_SPREAD = ( ORDER == 2 ) ? 1 : 0;

template <int... FORMS>
struct FormList
{
    static constexpr int N = sizeof...( FORMS );
    static constexpr int v[N] = { FORMS... };
};

constexpr int ng_of( int form )
{
    return ( form == FORM_SV || form == FORM_VD ) ? 1 : 3;
}
constexpr int np_of( int form )
{
    return ( form == FORM_SV || form == FORM_SG )
               ? 1
               : ( ( form == FORM_VV || form == FORM_VD ) ? 3 : 9 );
}
// LDS dof stride of a staged grid: 3-dof nodes are padded to 4 so one
// 128-bit LDS read fetches a node.
constexpr int ngp_of( int form ) { return ng_of( form ) == 1 ? 1 : 4; }

__device__ __forceinline__ size_t align16( size_t b )
{
    return ( b + 15 ) & ~(size_t)15;
}

struct TileGeom
{
    int tile;
    int c0[3];
    int c1[3];
};

__device__ __forceinline__ bool tile_geometry( const TileArgs& ta,
                                               const MeshDev& m, TileGeom& g )
{
    // Blocks are dealt round-robin to the 8 XCDs: give each XCD a
    // contiguous range of tiles so neighbouring windows share an L2.
    const int b = blockIdx.x;
    g.tile = ( b & 7 ) * ta.tiles_per_xcd + ( b >> 3 );
    const int ntiles = ta.ntile[0] * ta.ntile[1] * ta.ntile[2];
    if ( g.tile >= ntiles )
        return false;
    const int ti = g.tile / ( ta.ntile[1] * ta.ntile[2] );
    const int tj = ( g.tile / ta.ntile[2] ) % ta.ntile[1];
    const int tk = g.tile % ta.ntile[2];
    g.c0[0] = ti * ta.tile[0];
    g.c0[1] = tj * ta.tile[1];
    g.c0[2] = tk * ta.tile[2];
#pragma unroll
    for ( int d = 0; d < 3; ++d )
        g.c1[d] = min( g.c0[d] + ta.tile[d], m.num_cell[d] );
    return true;
}

// Fills the run table of the brick (one run per (i, j) cell row) and returns
// the number of particles.  run_pre has nruns + 1 entries.
__device__ __forceinline__ int setup_runs( const TileArgs& ta,
                                           const MeshDev& m,
                                           const TileGeom& g,
                                           int64_t* run_start, int* run_pre )
{
    const int nj = g.c1[1] - g.c0[1];
    const int nruns = ( g.c1[0] - g.c0[0] ) * nj;
    for ( int r = threadIdx.x; r < nruns; r += BLOCK )
    {
        const int ci = g.c0[0] + r / nj, cj = g.c0[1] + r % nj;
        const int64_t cfirst =
            ( (int64_t)ci * m.num_cell[1] + cj ) * m.num_cell[2] + g.c0[2];
        const int64_t clast = cfirst + ( g.c1[2] - g.c0[2] ) - 1;
        const int64_t st = ta.cell_offsets[cfirst];
        run_start[r] = st;
        run_pre[r + 1] = (int)( ta.cell_offsets[clast] + ta.cell_counts[clast] - st );
    }
    __syncthreads();
    if ( threadIdx.x == 0 )
    {
        run_pre[0] = 0;
        for ( int r = 0; r < nruns; ++r )
            run_pre[r + 1] += run_pre[r];
    }
    __syncthreads();
    return run_pre[nruns];
}

// Particle index of flattened brick position v (binary search of the runs).
__device__ __forceinline__ int64_t run_particle( const TileArgs& ta,
                                                 const int64_t* run_start,
                                                 const int* run_pre, int nruns,
                                                 int v )
{
    int lo = 0, hi = nruns - 1;
    while ( lo < hi )
    {
        const int mid = ( lo + hi + 1 ) >> 1;
        if ( run_pre[mid] <= v )
            lo = mid;
        else
            hi = mid - 1;
    }
    return ta.bin_begin + run_start[lo] + ( v - run_pre[lo] );
}

// Maps a flat window index to (node, dof) with the grid's unit-stride
// dimension fastest, so window loads/stores coalesce.
__device__ __forceinline__ void window_coords( const TileArgs& ta, int f,
                                               int ng, const int W[3], int q[3],
                                               int& d )
{
    const int pa = ta.flush_perm[0], pb = ta.flush_perm[1],
              pc = ta.flush_perm[2];
    d = f % ng;
    const int node = f / ng;
    q[pc] = node % W[pc];
    q[pb] = ( node / W[pc] ) % W[pb];
    q[pa] = node / ( W[pc] * W[pb] );
}

// ---------------------------------------------------------------------------
// Tiled gather
// ---------------------------------------------------------------------------

template <class Real, bool XBAR, bool SHARED, int... FORMS>
__global__ __launch_bounds__( BLOCK ) void gather_tiled_kernel(
    GatherArgs a, TileArgs ta )
{
    using FL = FormList<FORMS...>;
    constexpr int NT = FL::N;
    constexpr int NSLOT = SHARED ? 1 : NT;
    extern __shared__ __align__( 16 ) unsigned char smem[];

    TileGeom tg;
    if ( !tile_geometry( ta, a.mesh, tg ) )
        return;
    int wlo[3], W[3];
#pragma unroll
    for ( int d = 0; d < 3; ++d )
    {
        wlo[d] = tg.c0[d] + a.mesh.halo + BASE_SHIFT;
        W[d] = ta.win[d];
    }
    const int wnodes = W[0] * W[1] * W[2];

    // LDS: staged windows, then the run table.
    Real* win[NT];
    size_t off = 0;
    static_for<NT>( [&]( auto I ) {
        constexpr int t = I;
        if ( t < NSLOT )
        {
            win[t] = reinterpret_cast<Real*>( smem + off );
            off = align16( off + sizeof( Real ) * wnodes * ngp_of( FL::v[t] ) );
        }
        else
        {
            win[t] = win[0];
        }
    } );
    int64_t* run_start = reinterpret_cast<int64_t*>( smem + off );
    int* run_pre = reinterpret_cast<int*>( run_start + ta.max_runs );

    // Stage the windows (nodes outside the grid read as zero).
    static_for<NSLOT>( [&]( auto I ) {
        constexpr int t = I;
        constexpr int NG = ng_of( FL::v[t] );
        constexpr int NGP = ngp_of( FL::v[t] );
        const GFieldDev& g = a.t[t].grid;
        Real* w = win[t];
        for ( int f = threadIdx.x; f < wnodes * NG; f += BLOCK )
        {
            int q[3], d;
            window_coords( ta, f, NG, W, q, d );
            const int gi = wlo[0] + q[0], gj = wlo[1] + q[1],
                      gk = wlo[2] + q[2];
            Real v = Real( 0 );
            if ( gi >= 0 && gj >= 0 && gk >= 0 && gi < a.mesh.num_node[0] &&
                 gj < a.mesh.num_node[1] && gk < a.mesh.num_node[2] )
                v = *gptr<const Real>( g, gi * g.s[0] + gj * g.s[1] +
                                              gk * g.s[2] + d * g.s[3] );
            w[( ( q[0] * W[1] + q[1] ) * W[2] + q[2] ) * NGP + d] = v;
        }
    } );
    const int nruns = ( tg.c1[0] - tg.c0[0] ) * ( tg.c1[1] - tg.c0[1] );
    const int total = setup_runs( ta, a.mesh, tg, run_start, run_pre );

    for ( int v = threadIdx.x; v < total; v += BLOCK )
    {
        const int64_t p = run_particle( ta, run_start, run_pre, nruns, v );
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
            fits = fits && rel[d] >= 0 && rel[d] + K <= W[d];
        }
        if ( fits )
        {
            const int nb = ( rel[0] * W[1] + rel[1] ) * W[2] + rel[2];
            const int W1 = W[1], W2 = W[2];
            auto ld = [&]( auto S, int i, int j, int k, Real* f ) {
                constexpr int sl = decltype( S )::value;
                constexpr int NG = GatherNeeds<FORMS...>::NG( sl, SHARED );
                const int node = nb + ( i * W1 + j ) * W2 + k;
                if constexpr ( NG == 1 )
                {
                    f[0] = win[sl][node];
                }
                else
                {
                    const Real* q = win[sl] + node * 4;
                    f[0] = q[0];
                    f[1] = q[1];
                    f[2] = q[2];
                }
            };
            gather_particle<ORDER, Real, XBAR, SHARED, FORMS...>( a, p, s, ld );
        }
        else
        {
            // Stencil outside the staged window (particle not sorted into
            // this brick): read the grid directly.
            GlobalGatherLoad<Real, SHARED, FORMS...> ld{
                a, { s.d[0].base, s.d[1].base, s.d[2].base } };
            gather_particle<ORDER, Real, XBAR, SHARED, FORMS...>( a, p, s, ld );
        }
    }
}

// ---------------------------------------------------------------------------
// Cell-register scatter
// ---------------------------------------------------------------------------

// Exclusive scan of n (<= BLOCK * 16) ints in LDS, in place; returns the sum
// to every thread.
__device__ __forceinline__ int block_exclusive_scan( int* v, int n, int* wsum )
{
    const int per = ( n + BLOCK - 1 ) / BLOCK;
    const int lo = threadIdx.x * per;
    const int hi = min( lo + per, n );
    int s = 0;
    for ( int i = lo; i < hi; ++i )
        s += v[i];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int x = s;
#pragma unroll
    for ( int o = 1; o < 64; o <<= 1 )
    {
        const int y = __shfl_up( x, o, 64 );
        if ( lane >= o )
            x += y;
    }
    if ( lane == 63 )
        wsum[wv] = x;
    __syncthreads();
    int wbase = 0;
    for ( int w = 0; w < wv; ++w )
        wbase += wsum[w];
    const int total = wsum[0] + wsum[1] + wsum[2] + wsum[3];
    int run = wbase + x - s;
    for ( int i = lo; i < hi; ++i )
    {
        const int c = v[i];
        v[i] = run;
        run += c;
    }
    __syncthreads();
    return total;
}

template <class Real>
__device__ __forceinline__ Real pick( const Real* v, int q )
{
    Real r = v[0];
#pragma unroll
    for ( int i = 1; i < K; ++i )
        r = ( q == i ) ? v[i] : r;
    return r;
}

// Global-atomic scatter of one particle (particles whose stencil is outside
// the brick's window).
template <class Real, bool SHARED, int... FORMS>
__device__ __forceinline__ void global_scatter_particle( const ScatterArgs& a,
                                                      int64_t p,
                                                      const Spline3D<ORDER, Real, false>& s )
{
    using FL = FormList<FORMS...>;
    constexpr int NT = FL::N;
    Real Pm[NT][9];
    static_for<NT>( [&]( auto I ) {
        constexpr int t = I;
        constexpr int NP = np_of( FL::v[t] );
        const Real m = (Real)a.t[t].mult;
        const int64_t b = a.t[t].pin.base( p );
#pragma unroll
        for ( int c = 0; c < NP; ++c )
            Pm[t][c] = *a.t[t].pin.template ptr<const Real>( b, c ) * m;
    } );
#pragma unroll
    for ( int i = 0; i < K; ++i )
#pragma unroll
        for ( int j = 0; j < K; ++j )
#pragma unroll
            for ( int k = 0; k < K; ++k )
            {
                const auto n = node_weights<ORDER, Real, false, false>( s, i, j, k );
                static_for<NT>( [&]( auto I ) {
                    constexpr int t = I;
                    constexpr int FORM = FL::v[t];
                    constexpr int NG = ng_of( FORM );
                    const GFieldDev& g = SHARED ? a.t[0].grid : a.t[t].grid;
                    Real G[NG];
                    form_scatter<FORM>( Pm[t], n, G );
                    Real* q = gptr<Real>( g, ( s.d[0].base + i ) * g.s[0] +
                              ( s.d[1].base + j ) * g.s[1] +
                              ( s.d[2].base + k ) * g.s[2] );
#pragma unroll
                    for ( int d = 0; d < NG; ++d )
                        atomic_add( q + d * g.s[3], G[d] );
                } );
            }
}

template <int... FORMS>
struct SlotDofs
{
    using FL = FormList<FORMS...>;
    // Offset of term t's dofs in the packed accumulator (SHARED: all 0).
    static constexpr int offset( int t, bool shared )
    {
        int o = 0;
        for ( int i = 0; i < t; ++i )
            o += shared ? 0 : ng_of( FL::v[i] );
        return o;
    }
    static constexpr int total( bool shared )
    {
        int o = 0;
        for ( int i = 0; i < FL::N; ++i )
            o += ng_of( FL::v[i] );
        return shared ? 3 : o;
    }
};

template <class Real, bool SHARED, int... FORMS>
__global__ __launch_bounds__( BLOCK ) void scatter_cells_kernel(
    ScatterArgs a, TileArgs ta )
{
    using FL = FormList<FORMS...>;
    using SD = SlotDofs<FORMS...>;
    constexpr int NT = FL::N;
    constexpr int DT = SD::total( SHARED );
    extern __shared__ __align__( 16 ) unsigned char smem[];

    TileGeom tg;
    if ( !tile_geometry( ta, a.mesh, tg ) )
        return;
    // Base cells of the brick and the node window they reach.
    int blo[3], NB[3], W[3];
#pragma unroll
    for ( int d = 0; d < 3; ++d )
    {
        blo[d] = tg.c0[d] + a.mesh.halo + BASE_SHIFT;
        NB[d] = ( tg.c1[d] - tg.c0[d] ) + BASE_SPREAD;
        W[d] = ta.win[d];
    }
    const int nbt = NB[0] * NB[1] * NB[2];
    const int wnodes = W[0] * W[1] * W[2];

    // LDS layout (capacities fixed on the host from the unclipped tile):
    // FP64 window [node][DT], base-cell offsets, batch keys, sorted list,
    // run table, scan scratch.
    double* win = reinterpret_cast<double*>( smem );
    size_t off = align16( sizeof( double ) * (size_t)DT * ta.max_win_nodes );
    int* cnt = reinterpret_cast<int*>( smem + off );
    off = align16( off + sizeof( int ) * ( ta.max_base_cells + 1 ) );
    int* key = reinterpret_cast<int*>( smem + off );
    off = align16( off + sizeof( int ) * ta.cap );
    int* list = reinterpret_cast<int*>( smem + off );
    off = align16( off + sizeof( int ) * ta.cap );
    int64_t* run_start = reinterpret_cast<int64_t*>( smem + off );
    off = align16( off + sizeof( int64_t ) * ta.max_runs );
    int* run_pre = reinterpret_cast<int*>( smem + off );
    off = align16( off + sizeof( int ) * ( ta.max_runs + 1 ) );
    int* wsum = reinterpret_cast<int*>( smem + off );

    for ( int e = threadIdx.x; e < DT * wnodes; e += BLOCK )
        win[e] = 0.0;
    const int nruns = ( tg.c1[0] - tg.c0[0] ) * ( tg.c1[1] - tg.c0[1] );
    const int total = setup_runs( ta, a.mesh, tg, run_start, run_pre );

    for ( int bstart = 0; bstart < total; bstart += ta.cap )
    {
        const int nbatch = min( ta.cap, total - bstart );
        for ( int c = threadIdx.x; c <= nbt; c += BLOCK )
            cnt[c] = 0;
        __syncthreads();
        // A: stencil base of every particle of the batch, counted per base
        // cell (the slot is the particle's rank in its cell).
        for ( int i = threadIdx.x; i < nbatch; i += BLOCK )
        {
            const int64_t p =
                run_particle( ta, run_start, run_pre, nruns, bstart + i );
            double px[3];
            load_position( a.points, p, px );
            int bs[3];
            int k = -1;
            if ( !stencil_base<ORDER>( a.mesh, px, bs ) )
            {
                atomicAdd( a.skipped, 1ull );
            }
            else
            {
                int l[3];
                bool inside = true;
#pragma unroll
                for ( int d = 0; d < 3; ++d )
                {
                    l[d] = bs[d] - blo[d];
                    inside = inside && l[d] >= 0 && l[d] < NB[d];
                }
                if ( inside )
                {
                    const int cell = ( l[0] * NB[1] + l[1] ) * NB[2] + l[2];
                    const int slot = atomicAdd( cnt + cell, 1 );
                    k = ( cell << 16 ) | slot;
                }
                else
                {
                    // Not sorted into this brick: add straight to the grid.
                    Spline3D<ORDER, Real, false> s;
                    eval_spline( a.mesh, px, s );
                    global_scatter_particle<Real, SHARED, FORMS...>( a, p, s );
                }
            }
            key[i] = k;
        }
        __syncthreads();
        // B: base-cell offsets; cnt[nbt] = number of sorted particles.
        const int nsorted = block_exclusive_scan( cnt, nbt, wsum );
        if ( threadIdx.x == 0 )
            cnt[nbt] = nsorted;
        // C: particles (index relative to the binned range) listed by base
        // cell.
        for ( int i = threadIdx.x; i < nbatch; i += BLOCK )
        {
            const int k = key[i];
            if ( k >= 0 )
                list[cnt[k >> 16] + ( k & 0xffff )] = (int)(
                    run_particle( ta, run_start, run_pre, nruns, bstart + i ) -
                    ta.bin_begin );
        }
        __syncthreads();
        // D: one lane per (base cell, x-slab): accumulate the slab's K^2
        // nodes over the cell's particles in registers, flush once.
        for ( int task = threadIdx.x; task < nbt * K; task += BLOCK )
        {
            const int cell = task / K;
            const int sa = task - cell * K;
            const int q0 = cnt[cell], q1 = cnt[cell + 1];
            if ( q0 == q1 )
                continue;
            Real acc[K * K][DT];
#pragma unroll
            for ( int n = 0; n < K * K; ++n )
#pragma unroll
                for ( int d = 0; d < DT; ++d )
                    acc[n][d] = Real( 0 );
            for ( int q = q0; q < q1; ++q )
            {
                const int64_t p = ta.bin_begin + list[q];
                double px[3];
                load_position( a.points, p, px );
                Spline1D<ORDER, Real> sx, sy, sz;
                eval_spline_1d<ORDER, Real, false>(
                    logical_coord<ORDER>( px[0], a.mesh.low[0], a.mesh.rdx[0] ),
                    (Real)a.mesh.rdx[0], a.mesh.num_node[0], sx );
                eval_spline_1d<ORDER, Real, false>(
                    logical_coord<ORDER>( px[1], a.mesh.low[1], a.mesh.rdx[1] ),
                    (Real)a.mesh.rdx[1], a.mesh.num_node[1], sy );
                eval_spline_1d<ORDER, Real, false>(
                    logical_coord<ORDER>( px[2], a.mesh.low[2], a.mesh.rdx[2] ),
                    (Real)a.mesh.rdx[2], a.mesh.num_node[2], sz );
                const Real wx = pick( sx.w, sa );
                const Real gx = pick( sx.g, sa );
                Real Pm[NT][9];
                static_for<NT>( [&]( auto I ) {
                    constexpr int t = I;
                    constexpr int NP = np_of( FL::v[t] );
                    const Real m = (Real)a.t[t].mult;
                    const int64_t b = a.t[t].pin.base( p );
#pragma unroll
                    for ( int c = 0; c < NP; ++c )
                        Pm[t][c] = *a.t[t].pin.template ptr<const Real>( b, c ) * m;
                } );
#pragma unroll
                for ( int j = 0; j < K; ++j )
#pragma unroll
                    for ( int k = 0; k < K; ++k )
                    {
                        NodeW<Real, false> n;
                        const Real wy = sy.w[j], wz = sz.w[k];
                        const Real wyz = wy * wz;
                        n.W = wx * wyz;
                        n.dW[0] = gx * wyz;
                        n.dW[1] = wx * sy.g[j] * wz;
                        n.dW[2] = wx * wy * sz.g[k];
                        static_for<NT>( [&]( auto I ) {
                            constexpr int t = I;
                            constexpr int FORM = FL::v[t];
                            constexpr int NG = ng_of( FORM );
                            constexpr int O = SD::offset( t, SHARED );
                            Real G[NG];
                            form_scatter<FORM>( Pm[t], n, G );
#pragma unroll
                            for ( int d = 0; d < NG; ++d )
                                acc[j * K + k][O + d] += G[d];
                        } );
                    }
            }
            const int lx = cell / ( NB[1] * NB[2] );
            const int ly = ( cell / NB[2] ) % NB[1];
            const int lz = cell % NB[2];
#pragma unroll
            for ( int j = 0; j < K; ++j )
#pragma unroll
                for ( int k = 0; k < K; ++k )
                {
                    const int node = ( ( lx + sa ) * W[1] + ly + j ) * W[2] + lz + k;
#pragma unroll
                    for ( int d = 0; d < DT; ++d )
                        atomicAdd( win + node * DT + d, (double)acc[j * K + k][d] );
                }
        }
        __syncthreads();
    }

    // Flush the window: one global atomic per non-zero node value, walking the
    // window with the grid's unit-stride dimension fastest.
    static_for<( SHARED ? 1 : NT )>( [&]( auto I ) {
        constexpr int t = I;
        constexpr int NG = SHARED ? 3 : ng_of( FL::v[t] );
        constexpr int O = SD::offset( t, SHARED );
        const GFieldDev& g = a.t[t].grid;
        for ( int f = threadIdx.x; f < wnodes * NG; f += BLOCK )
        {
            int q[3], d;
            window_coords( ta, f, NG, W, q, d );
            const Real v =
                (Real)win[( ( q[0] * W[1] + q[1] ) * W[2] + q[2] ) * DT + O + d];
            if ( v == Real( 0 ) )
                continue;
            const int gi = blo[0] + q[0], gj = blo[1] + q[1], gk = blo[2] + q[2];
            if ( gi < 0 || gj < 0 || gk < 0 || gi >= a.mesh.num_node[0] ||
                 gj >= a.mesh.num_node[1] || gk >= a.mesh.num_node[2] )
                continue;
            atomic_add( gptr<Real>( g, gi * g.s[0] +
                            gj * g.s[1] + gk * g.s[2] + d * g.s[3] ),
                        v );
        }
    } );
}

// ---------------------------------------------------------------------------
// Launch helpers
// ---------------------------------------------------------------------------

template <class KernelPtr, class Args>
hipError_t launch_tile_kernel( KernelPtr k, const Args& a, const TileArgs& t,
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
    hipLaunchKernelGGL( k, dim3( (unsigned)( 8 * t.tiles_per_xcd ) ),
                        dim3( BLOCK ), lds, s, a, t );
    return hipGetLastError();
}

template <class Real>
hipError_t gather_tiled_typed( Launcher& L, int plan, const GatherArgs& a,
                               const TileArgs& t, size_t lds, hipStream_t s )
{
    const char* name =
        a.has_xbar ? ( plan == PLAN_MPM_P2G ? "p2g_adjoint_mpm_tiled" : "gather_xbar_tiled" )
                   : ( plan == PLAN_MPM_G2P ? "g2p_mpm_tiled" : "g2p_tiled" );
    hipError_t e = hipSuccess;
    L.before( name, s );
    if ( plan == PLAN_SINGLE )
    {
        switch ( a.t[0].form )
        {
#define CA_CASE( F )                                                           \
    case F:                                                                    \
        e = a.has_xbar ? launch_tile_kernel( gather_tiled_kernel<Real, true, false, F>, \
                                             a, t, lds, s )                    \
                       : launch_tile_kernel( gather_tiled_kernel<Real, false, false, F>, \
                                             a, t, lds, s );                   \
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
    else if ( plan == PLAN_MPM_G2P && !a.has_xbar )
        e = launch_tile_kernel( gather_tiled_kernel<Real, false, true, FORM_VV, FORM_TDT>,
                                a, t, lds, s );
    else if ( plan == PLAN_MPM_G2P && a.has_xbar )
        e = launch_tile_kernel( gather_tiled_kernel<Real, true, true, FORM_VV, FORM_TDT>,
                                a, t, lds, s );
    else if ( plan == PLAN_MPM_P2G && a.has_xbar )
        e = launch_tile_kernel(
            gather_tiled_kernel<Real, true, false, FORM_SV, FORM_VV, FORM_TD>, a, t,
            lds, s );
    else
        e = hipErrorInvalidValue;
    L.after( name, s );
    return e;
}

template <class Real>
hipError_t scatter_cells_typed( Launcher& L, int plan, const ScatterArgs& a,
                                const TileArgs& t, size_t lds, hipStream_t s )
{
    const char* name = plan == PLAN_MPM_P2G   ? "p2g_mpm_cells"
                       : plan == PLAN_MPM_G2P ? "g2p_adjoint_mpm_cells"
                                              : "scatter_cells";
    hipError_t e = hipSuccess;
    L.before( name, s );
    if ( plan == PLAN_SINGLE )
    {
        switch ( a.t[0].form )
        {
#define CA_CASE( F )                                                           \
    case F:                                                                    \
        e = launch_tile_kernel( scatter_cells_kernel<Real, false, F>, a, t, lds, s ); \
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
    else if ( plan == PLAN_MPM_P2G )
        e = launch_tile_kernel(
            scatter_cells_kernel<Real, false, FORM_SV, FORM_VV, FORM_TD>, a, t, lds, s );
    else if ( plan == PLAN_MPM_G2P )
        e = launch_tile_kernel( scatter_cells_kernel<Real, true, FORM_VV, FORM_TDT>, a,
                                t, lds, s );
    else
        e = hipErrorInvalidValue;
    L.after( name, s );
    return e;
}

} // namespace

hipError_t CA_CAT( launch_gather_tiled_, CA_SUFFIX )( Launcher& L, bool fp64,
                                                     int plan,
                                                     const GatherArgs& a,
                                                     const TileArgs& t,
                                                     size_t lds, hipStream_t s )
{
    return fp64 ? gather_tiled_typed<double>( L, plan, a, t, lds, s )
                : gather_tiled_typed<float>( L, plan, a, t, lds, s );
}

hipError_t CA_CAT( launch_scatter_cells_, CA_SUFFIX )( Launcher& L, bool fp64,
                                                      int plan,
                                                      const ScatterArgs& a,
                                                      const TileArgs& t,
                                                      size_t lds, hipStream_t s )
{
    return fp64 ? scatter_cells_typed<double>( L, plan, a, t, lds, s )
                : scatter_cells_typed<float>( L, plan, a, t, lds, s );
}

} // namespace cabana_amd
