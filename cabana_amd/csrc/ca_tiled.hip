// Cell-tiled interpolation kernels for cell-sorted particles, one spline
// order per translation unit (CA_ORDER set on the compile line).
//
// Bricks of mesh cells are cut into parts of bounded particle count (a part
// table built on the device before the launch), and each workgroup walks
// parts from an XCD-local range of that table, so dense clusters are spread
// over every CU and empty bricks cost nothing.  A brick's particles are its
// cell runs along the last dimension (contiguous in the cell-sorted order),
// walked through a flattened index so consecutive lanes read consecutive
// particles.
//
//   gather_tiled_kernel   G2P forward, P2G adjoint and the position part of
//                         the G2P adjoint.  The node window of every grid the
//                         pass reads is staged in LDS with coalesced loads;
//                         every particle then gathers its K^3 stencil from
//                         LDS.
//   scatter_cells_kernel  P2G forward and the grid part of the G2P adjoint.
//                         The part's particles are counting-sorted in LDS by
//                         stencil base cell (records: FP32/FP64 cell offsets
//                         + premultiplied attributes, loaded with coalesced
//                         reads a part ahead).  Each base cell's records form
//                         tasks of at most SCATTER_TASK_MAX, ordered longest
//                         first; a task's K*K lanes each own one (j, k) row of
//                         the stencil, accumulate its K nodes for every term
//                         in registers, and add them once into an FP64 LDS
//                         window (LDS FP64 adds are ~16x faster than FP32 ones
//                         on gfx950).  The window is flushed into the grid
//                         once per part.
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
// Scatter record header.  FP32 records carry the x weights and weight
// gradients of the K nodes, then the y and z fractional offsets (the x
// weights are formed once per particle instead of by each of a task's K*K
// lanes); FP64 records, whose header would double in bytes, carry the
// three fractional offsets.
template <class Real>
constexpr bool x_weights_in_record()
{
    return sizeof( Real ) == 4;
}
template <class Real>
constexpr int rhead()
{
    return x_weights_in_record<Real>() ? 2 * K + 2 : 3;
}

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
    int brick;
    int part;
    int c0[3];
    int c1[3];
};

// The parts of this workgroup: the part table is split into 8 contiguous
// ranges, one per XCD (workgroups are dealt round-robin to the XCDs), so
// neighbouring bricks share an L2; the workgroups of one XCD stride through
// its range.  Every workgroup leaves the loop once its range is exhausted.
struct PartLoop
{
    int total;
    int per_xcd;
    int xcd;
    int first;
    int stride;
    __device__ explicit PartLoop( const TileArgs& ta )
    {
        total = *ta.num_parts;
        per_xcd = ( total + 7 ) >> 3;
        xcd = blockIdx.x & 7;
        first = blockIdx.x >> 3;
        stride = gridDim.x >> 3;
    }
};

__device__ __forceinline__ void part_geometry_of( const TileArgs& ta, const MeshDev& m,
                                                  int2 pr, TileGeom& g )
{
    g.brick = pr.x;
    g.part = pr.y;
    const int ti = g.brick / ( ta.ntile[1] * ta.ntile[2] );
    const int tj = ( g.brick / ta.ntile[2] ) % ta.ntile[1];
    const int tk = g.brick % ta.ntile[2];
    g.c0[0] = ti * ta.tile[0];
    g.c0[1] = tj * ta.tile[1];
    g.c0[2] = tk * ta.tile[2];
#pragma unroll
    for ( int d = 0; d < 3; ++d )
        g.c1[d] = min( g.c0[d] + ta.tile[d], m.num_cell[d] );
}

__device__ __forceinline__ void part_geometry( const TileArgs& ta, const MeshDev& m,
                                               int w, TileGeom& g )
{
    part_geometry_of( ta, m, ta.parts[w], g );
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
    return ta.particle( ta.bin_begin + run_start[lo] + ( v - run_pre[lo] ) );
}

// Run containing flattened brick position v (binary search).
__device__ __forceinline__ int find_run( const int* run_pre, int nruns, int v )
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
    return lo;
}

// floor(a / b) for 0 <= a < 2^22 and b >= 1 from a rounded reciprocal rb of
// b: (a + 1/2) / b lies at least 1/(2b) from an integer, and the two
// roundings move it by at most (a + 1/2) / b * 2^-23 < 1/(2b).
__device__ __forceinline__ int div_small( int a, float rb )
{
    return (int)( ( (float)a + 0.5f ) * rb );
}

// Maps a flat window index to (node, dof) with the grid's unit-stride
// dimension fastest, so window loads/stores coalesce.  rW holds the
// reciprocals of the window extents (window indices stay below 2^22).
template <int NG>
__device__ __forceinline__ void window_coords( const TileArgs& ta, int f, const int W[3],
                                               const float rW[3], int q[3], int& d )
{
    const int pa = ta.flush_perm[0], pb = ta.flush_perm[1],
              pc = ta.flush_perm[2];
    const int node = f / NG;
    d = f - node * NG;
    const int r = div_small( node, rW[pc] );
    q[pc] = node - r * W[pc];
    q[pa] = div_small( r, rW[pb] );
    q[pb] = r - q[pa] * W[pb];
}

// ---------------------------------------------------------------------------
// Tiled gather
// ---------------------------------------------------------------------------

// Waves per SIMD the register allocation of a gather must leave room for:
// three for the FP32 passes that read their inputs late (more than three
// terms; their LDS windows allow four), no bound otherwise.
template <class Real, int... FORMS>
constexpr int gather_min_waves()
{
    return ( sizeof( Real ) == 4 && GatherNeeds<FORMS...>::late() ) ? 3 : 1;
}

template <class Real, int MODE, bool CONTIG, bool SHARED, int... FORMS>
__global__ __launch_bounds__( BLOCK, ( gather_min_waves<Real, FORMS...>() ) ) void gather_tiled_kernel(
    GatherArgs a, TileArgs ta )
{
    using GN = GatherNeeds<FORMS...>;
    constexpr int NSLOT = GN::nslot( SHARED );
    extern __shared__ __align__( 16 ) unsigned char smem[];

    // LDS: one staged window per grid slot (3-dof slots padded to 4), then
    // the run table (capacities of a full brick).
    Real* win[NSLOT];
    size_t off = 0;
    static_for<NSLOT>( [&]( auto I ) {
        constexpr int sl = I;
        win[sl] = reinterpret_cast<Real*>( smem + off );
        off = align16( off + sizeof( Real ) * ta.max_win_nodes *
                                 ( GN::NG( sl, SHARED ) == 1 ? 1 : 4 ) );
    } );
    int64_t* run_start = reinterpret_cast<int64_t*>( smem + off );
    int* run_pre = reinterpret_cast<int*>( run_start + ta.max_runs );

    const PartLoop pl( ta );
    for ( int loc = pl.first; loc < pl.per_xcd; loc += pl.stride )
    {
        const int w = pl.xcd * pl.per_xcd + loc;
        if ( w >= pl.total )
            break;
        TileGeom tg;
        part_geometry( ta, a.mesh, w, tg );
        int wlo[3], W[3];
#pragma unroll
        for ( int d = 0; d < 3; ++d )
        {
            wlo[d] = tg.c0[d] + a.mesh.halo + BASE_SHIFT;
            W[d] = ta.win[d];
        }
        const int wnodes = W[0] * W[1] * W[2];
        const float rW[3] = { 1.0f / (float)W[0], 1.0f / (float)W[1], 1.0f / (float)W[2] };

        // Stage the windows (nodes outside the grid read as zero).
        static_for<NSLOT>( [&]( auto I ) {
            constexpr int sl = I;
            constexpr int NG = GN::NG( sl, SHARED );
            constexpr int NGP = NG == 1 ? 1 : 4;
            const GFieldDev& g = a.t[GN::first_term( sl, SHARED )].grid;
            Real* wp = win[sl];
            for ( int f = threadIdx.x; f < wnodes * NG; f += BLOCK )
            {
                int q[3], d;
                window_coords<NG>( ta, f, W, rW, q, d );
                const int gi = wlo[0] + q[0], gj = wlo[1] + q[1], gk = wlo[2] + q[2];
                Real v = Real( 0 );
                if ( gi >= 0 && gj >= 0 && gk >= 0 && gi < a.mesh.num_node[0] &&
                     gj < a.mesh.num_node[1] && gk < a.mesh.num_node[2] )
                    v = *gptr<const Real>( g, gi * g.s[0] + gj * g.s[1] + gk * g.s[2] +
                                                  d * g.s[3] );
                wp[( ( q[0] * W[1] + q[1] ) * W[2] + q[2] ) * NGP + d] = v;
            }
        } );
        const int nruns = ( tg.c1[0] - tg.c0[0] ) * ( tg.c1[1] - tg.c0[1] );
        const int total = setup_runs( ta, a.mesh, tg, run_start, run_pre );
        const int pbeg = tg.part * ta.cap;
        const int pend = min( total, pbeg + ta.cap );

        // Lanes stride through the part; the cell run of the next particle
        // is found by walking forward, and its position and other global
        // reads are issued one iteration ahead, so their latency overlaps
        // the current particle's stencil work.
        using In = GatherIn<Real, MODE, FORMS...>;
        int v = pbeg + (int)threadIdx.x;
        int r = 0;
        int64_t p = 0;
        double px[3] = { 0.0, 0.0, 0.0 };
        In in;
        if ( v < pend )
        {
            r = find_run( run_pre, nruns, v );
            p = ta.particle( ta.bin_begin + run_start[r] + ( v - run_pre[r] ) );
            load_position<CONTIG>( a.points, p, px );
            gather_load<Real, MODE, CONTIG, FORMS...>( a, p, in );
        }
        while ( v < pend )
        {
            const int vn = v + BLOCK;
            int64_t pn = 0;
            double pxn[3] = { 0.0, 0.0, 0.0 };
            In inn;
            if ( vn < pend )
            {
                while ( run_pre[r + 1] <= vn )
                    ++r;
                pn = ta.particle( ta.bin_begin + run_start[r] + ( vn - run_pre[r] ) );
                load_position<CONTIG>( a.points, pn, pxn );
                gather_load<Real, MODE, CONTIG, FORMS...>( a, pn, inn );
            }
            Spline3D<ORDER, Real, mode_xbar( MODE )> s;
            if ( !eval_spline( a.mesh, px, s ) )
            {
                atomicAdd( a.skipped, 1ull );
            }
            else
            {
                int rel[3];
                bool fits = true;
#pragma unroll
                for ( int d = 0; d < 3; ++d )
                {
                    rel[d] = s.d[d].base - wlo[d];
                    fits = fits && rel[d] >= 0 && rel[d] + K <= W[d];
                }
                // Stencil nodes from the staged window.  A wave with a
                // particle whose stencil is outside the window (not sorted
                // into this brick) reads the grid directly for all of its
                // lanes, so the choice is one uniform branch, not one per
                // node.
                if ( __all( fits ) )
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
                    gather_compute<ORDER, Real, MODE, CONTIG, SHARED, FORMS...>( a, p, s, ld, in );
                }
                else
                {
                    GlobalGatherLoad<Real, SHARED, FORMS...> ld{
                        a, { s.d[0].base, s.d[1].base, s.d[2].base } };
                    gather_compute<ORDER, Real, MODE, CONTIG, SHARED, FORMS...>( a, p, s, ld, in );
                }
            }
            v = vn;
            p = pn;
#pragma unroll
            for ( int d = 0; d < 3; ++d )
                px[d] = pxn[d];
            in = inn;
        }
        // The next part restages the windows.
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Cell-register scatter
// ---------------------------------------------------------------------------

// Exclusive scan of n (<= NTH * 16) ints in LDS, in place, by a workgroup of
// NTH threads; returns the sum to every thread.
struct ScanIdentity
{
    __device__ int operator()( int x ) const { return x; }
};

// Exclusive scan of f(v[i]) (f applied as the values are read).
template <int NTH, class F = ScanIdentity>
__device__ __forceinline__ int block_exclusive_scan( int* v, int n, int* wsum, F f = F() )
{
    const int per = ( n + NTH - 1 ) / NTH;
    const int lo = threadIdx.x * per;
    const int hi = min( lo + per, n );
    int s = 0;
    for ( int i = lo; i < hi; ++i )
    {
        v[i] = f( v[i] );
        s += v[i];
    }
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
    int wbase = 0, total = 0;
#pragma unroll
    for ( int w = 0; w < NTH / 64; ++w )
    {
        const int c = wsum[w];
        wbase += w < wv ? c : 0;
        total += c;
    }
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
        const uint32_t b = a.t[t].pin.base( p );
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

__device__ __forceinline__ float fma_r( float a, float b, float c ) { return fmaf( a, b, c ); }
__device__ __forceinline__ double fma_r( double a, double b, double c ) { return fma( a, b, c ); }


// Record layout of a scatter pass: the rhead<Real>() header, then the
// attributes of every term in order.
template <class Real, int... FORMS>
struct RecordLayout
{
    using FL = FormList<FORMS...>;
    static constexpr int attr( int t )
    {
        int o = rhead<Real>();
        for ( int i = 0; i < t; ++i )
            o += np_of( FL::v[i] );
        return o;
    }
    static constexpr int len() { return attr( FL::N ); }
};

// Task-table entry: base cell (x, y, z: 4 bits each), first record (15
// bits), count - 1 (5 bits: 1 to SCATTER_TASK_MAX records).
static_assert( SCATTER_TASK_MAX <= 32, "task entries hold count - 1 in 5 bits" );
__device__ __forceinline__ unsigned pack_task( int lx, int ly, int lz, int start, int len )
{
    return ( (unsigned)lx << 28 ) | ( (unsigned)ly << 24 ) | ( (unsigned)lz << 20 ) |
           ( (unsigned)start << 5 ) | (unsigned)( len - 1 );
}

// The run table of a brick in two halves, so its global reads can be issued
// early and committed later: lane r < nruns (<= 64) of wave 0 holds run r's
// raw reads (nothing is computed from them until the commit, so issuing
// them does not wait for them).
struct RunLoad
{
    int64_t start;
    int64_t last_offset;
    int32_t last_count;
};

__device__ __forceinline__ void runs_issue( const TileArgs& ta, const MeshDev& m,
                                            const TileGeom& g, RunLoad& rl )
{
    const int nj = g.c1[1] - g.c0[1];
    const int nruns = ( g.c1[0] - g.c0[0] ) * nj;
    const int r = threadIdx.x;
    rl.start = 0;
    rl.last_offset = 0;
    rl.last_count = 0;
    if ( r < nruns )
    {
        const int ci = g.c0[0] + r / nj, cj = g.c0[1] + r % nj;
        const int64_t cfirst = ( (int64_t)ci * m.num_cell[1] + cj ) * m.num_cell[2] + g.c0[2];
        const int64_t clast = cfirst + ( g.c1[2] - g.c0[2] ) - 1;
        rl.start = ta.cell_offsets[cfirst];
        rl.last_offset = ta.cell_offsets[clast];
        rl.last_count = ta.cell_counts[clast];
    }
}

// Wave 0 writes run_start and the inclusive prefix run_pre (visible after
// the next barrier).
__device__ __forceinline__ void runs_commit( const RunLoad& rl, const TileGeom& g,
                                             int64_t* run_start, int* run_pre )
{
    if ( threadIdx.x >= 64 )
        return;
    const int nruns = ( g.c1[0] - g.c0[0] ) * ( g.c1[1] - g.c0[1] );
    const int lane = threadIdx.x;
    int v = (int)( rl.last_offset + rl.last_count - rl.start );
#pragma unroll
    for ( int o = 1; o < 64; o <<= 1 )
    {
        const int u = __shfl_up( v, o, 64 );
        if ( lane >= o )
            v += u;
    }
    if ( lane < nruns )
    {
        run_start[lane] = rl.start;
        run_pre[lane + 1] = v;
    }
    if ( lane == 0 )
        run_pre[0] = 0;
}

// One stencil node's weight polynomial (spline_coef / spline_shift of the
// lane's node index), with the derivative's coefficients pre-scaled by rdx,
// held per lane.
template <class Real>
struct NodePoly
{
    Real shift;
    Real w[ORDER + 1];
    Real g[ORDER];
    __device__ __forceinline__ void init( int j, Real rdx )
    {
        // Node j's shift and coefficients, selected from compile-time
        // constants.
        static_for<K>( [&]( auto JJ ) {
            constexpr int jj = JJ;
            constexpr double sh = spline_shift( ORDER, jj );
            shift = ( jj == 0 || j == jj ) ? Real( sh ) : shift;
            static_for<ORDER + 1>( [&]( auto N ) {
                constexpr int n = N;
                constexpr double c = spline_coef( ORDER, jj, n );
                w[n] = ( jj == 0 || j == jj ) ? Real( c ) : w[n];
            } );
        } );
#pragma unroll
        for ( int n = 1; n <= ORDER; ++n )
            g[n - 1] = Real( n ) * w[n] * rdx;
    }
    __device__ __forceinline__ void eval( Real f, Real& wv, Real& gv ) const
    {
        const Real t = f - shift;
        wv = w[ORDER];
#pragma unroll
        for ( int n = ORDER - 1; n >= 0; --n )
            wv = fma_r( wv, t, w[n] );
        gv = g[ORDER - 1];
#pragma unroll
        for ( int n = ORDER - 2; n >= 0; --n )
            gv = fma_r( gv, t, g[n] );
    }
};

// Pencil form of a node value.  With A = wy wz, B = gy wz and C = wy gz of
// a fixed (j, k) node row, the node at stencil index i along x has weight
// wx_i A and weight gradient (gx_i A, wx_i B, wx_i C).  Every form's node
// value (dB/dG of the forms in ca_device.hpp) is then
//   G[d] = gx_i * alpha[d] + wx_i * beta[d]
// with alpha and beta formed once per particle and row.  Value-type dofs
// have no alpha; the y/z-gradient dofs of FORM_SG have no alpha either, and
// its x-gradient dof no beta.
template <int FORM>
__host__ __device__ constexpr bool form_has_alpha( int d )
{
    return FORM == FORM_SG ? d == 0 : !( FORM == FORM_SV || FORM == FORM_VV );
}
template <int FORM>
__host__ __device__ constexpr bool form_has_beta( int d )
{
    return FORM == FORM_SG ? d != 0 : true;
}

template <int FORM, class Real>
__device__ __forceinline__ void pencil_coeffs( const Real* P, Real A, Real B, Real C,
                                               Real* al, Real* be )
{
    if constexpr ( FORM == FORM_SV )
    {
        be[0] = fma_r( A, P[0], be[0] );
    }
    else if constexpr ( FORM == FORM_VV )
    {
#pragma unroll
        for ( int d = 0; d < 3; ++d )
            be[d] = fma_r( A, P[d], be[d] );
    }
    else if constexpr ( FORM == FORM_SG )
    {
        al[0] = fma_r( A, P[0], al[0] );
        be[1] = fma_r( B, P[0], be[1] );
        be[2] = fma_r( C, P[0], be[2] );
    }
    else if constexpr ( FORM == FORM_VD )
    {
        al[0] = fma_r( A, P[0], al[0] );
        be[0] = fma_r( B, P[1], fma_r( C, P[2], be[0] ) );
    }
    else if constexpr ( FORM == FORM_TD )
    {
#pragma unroll
        for ( int a = 0; a < 3; ++a )
        {
            al[a] = fma_r( A, P[3 * a], al[a] );
            be[a] = fma_r( B, P[3 * a + 1], fma_r( C, P[3 * a + 2], be[a] ) );
        }
    }
    else
    {
#pragma unroll
        for ( int b = 0; b < 3; ++b )
        {
            al[b] = fma_r( A, P[b], al[b] );
            be[b] = fma_r( B, P[3 + b], fma_r( C, P[6 + b], be[b] ) );
        }
    }
}

// Accumulator dofs of a scatter pass that receive an alpha / beta term.
template <bool SHARED, int... FORMS>
struct PencilDofs
{
    using FL = FormList<FORMS...>;
    using SD = SlotDofs<FORMS...>;
    template <int FORM>
    static constexpr bool pick( int d, bool alpha )
    {
        return alpha ? form_has_alpha<FORM>( d ) : form_has_beta<FORM>( d );
    }
    static constexpr bool has( int d, bool alpha )
    {
        bool r = false;
        int o = 0;
        ( ( r = r || ( d >= ( SHARED ? 0 : o ) && d < ( SHARED ? 0 : o ) + ng_of( FORMS ) &&
                       pick<FORMS>( d - ( SHARED ? 0 : o ), alpha ) ),
            o += ng_of( FORMS ) ),
          ... );
        return r;
    }
};

// One record's contribution to a lane's K x DT accumulator (nodes i of the
// lane's (j, k) row).  The record holds the particle's header (rhead) and
// its premultiplied attributes.
template <class Real, bool SHARED, int... FORMS>
__device__ __forceinline__ void pencil_record( const Real* rp, Real rdx0, const NodePoly<Real>& py,
                                               const NodePoly<Real>& pz,
                                               Real ( *acc )[SlotDofs<FORMS...>::total( SHARED )] )
{
    using FL = FormList<FORMS...>;
    using SD = SlotDofs<FORMS...>;
    using RL = RecordLayout<Real, FORMS...>;
    using PD = PencilDofs<SHARED, FORMS...>;
    constexpr int DT = SD::total( SHARED );
    const Real* wx;
    const Real* gx;
    Real wy, gy, wz, gz;
    Spline1D<ORDER, Real> sx;
    if constexpr ( x_weights_in_record<Real>() )
    {
        wx = rp;
        gx = rp + K;
        py.eval( rp[2 * K], wy, gy );
        pz.eval( rp[2 * K + 1], wz, gz );
    }
    else
    {
        spline_weights<ORDER, Real, false>( rp[0], rdx0, sx );
        wx = sx.w;
        gx = sx.g;
        py.eval( rp[1], wy, gy );
        pz.eval( rp[2], wz, gz );
    }
    const Real A = wy * wz, B = gy * wz, C = wy * gz;
    Real al[DT], be[DT];
#pragma unroll
    for ( int d = 0; d < DT; ++d )
        al[d] = be[d] = Real( 0 );
    static_for<FL::N>( [&]( auto I ) {
        constexpr int t = I;
        constexpr int O = SD::offset( t, SHARED );
        pencil_coeffs<FL::v[t]>( rp + RL::attr( t ), A, B, C, al + O, be + O );
    } );
#pragma unroll
    for ( int i = 0; i < K; ++i )
        static_for<DT>( [&]( auto D ) {
            constexpr int d = D;
            if constexpr ( PD::has( d, true ) )
                acc[i][d] = fma_r( gx[i], al[d], acc[i][d] );
            if constexpr ( PD::has( d, false ) )
                acc[i][d] = fma_r( wx[i], be[d], acc[i][d] );
        } );
}

// Reads one LDS record (16-byte aligned) into registers.
template <class Real, int N>
__device__ __forceinline__ void read_record( const Real* rp, Real* v )
{
    constexpr int W = 16 / (int)sizeof( Real );
    typedef Real V __attribute__( ( ext_vector_type( W ) ) );
#pragma unroll
    for ( int c = 0; c < N; c += W )
    {
        const V x = *reinterpret_cast<const V*>( rp + c );
#pragma unroll
        for ( int e = 0; e < W; ++e )
            if ( c + e < N )
                v[c + e] = x[e];
    }
}

// Particles of one part held in registers between their loads and their use
// (the loads of part k + 1 are in flight while part k accumulates).
template <class Real, int MAXM, int NPT>
struct PartPrefetch
{
    int p[MAXM];
    double x[MAXM][3];
    Real attr[MAXM][NPT];
};

template <class Real, bool CONTIG, int MAXM, int NTH, int... FORMS>
__device__ __forceinline__ void prefetch_part( const ScatterArgs& a, const TileArgs& ta,
                                               const int64_t* run_start, const int* run_pre,
                                               int nruns, int pbeg, int nbatch,
                                               PartPrefetch<Real, MAXM, RecordLayout<Real, FORMS...>::len() - rhead<Real>()>& pf )
{
    using FL = FormList<FORMS...>;
    using RL = RecordLayout<Real, FORMS...>;
    // Run of every slot: count the run starts at or below its flattened
    // index (broadcast LDS reads, no dependent search).
    int run[MAXM];
#pragma unroll
    for ( int mi = 0; mi < MAXM; ++mi )
        run[mi] = 0;
    for ( int q = 1; q < nruns; ++q )
    {
        const int rp = run_pre[q];
#pragma unroll
        for ( int mi = 0; mi < MAXM; ++mi )
            run[mi] += ( rp <= pbeg + (int)threadIdx.x + mi * NTH ) ? 1 : 0;
    }
#pragma unroll
    for ( int mi = 0; mi < MAXM; ++mi )
    {
        const int i = (int)threadIdx.x + mi * NTH;
        pf.p[mi] = -1;
        if ( i >= nbatch )
            continue;
        const int64_t p =
            ta.particle( ta.bin_begin + run_start[run[mi]] + ( pbeg + i - run_pre[run[mi]] ) );
        pf.p[mi] = (int)p;
        load_position<CONTIG>( a.points, p, pf.x[mi] );
        static_for<FL::N>( [&]( auto I ) {
            constexpr int t = I;
            a.t[t].pin.template load<Real, np_of( FL::v[t] ), CONTIG>(
                p, &pf.attr[mi][RL::attr( t ) - rhead<Real>()] );
        } );
    }
}

// Workgroup size of the scatter and its minimum waves per SIMD (the
// register budget): two 256-thread workgroups per CU.  For FP32 records a
// bound of three fits the LDS too, but its 168 registers make the MPM pass
// spill (it needs about 190), which measured 8 % slower.
template <class Real>
constexpr int sblock()
{
    return scatter_block( sizeof( Real ) == 8 );
}
template <class Real>
constexpr int scatter_min_waves()
{
    return 2;
}

template <class Real, bool CONTIG, bool SHARED, int... FORMS>
__global__ __launch_bounds__( sblock<Real>(), scatter_min_waves<Real>() ) void scatter_cells_kernel(
    ScatterArgs a, TileArgs ta )
{
    using FL = FormList<FORMS...>;
    using SD = SlotDofs<FORMS...>;
    using RL = RecordLayout<Real, FORMS...>;
    constexpr int NT = FL::N;
    constexpr int DT = SD::total( SHARED );
    constexpr int RLEN = RL::len();
    constexpr int RS = record_stride( RLEN, (int)sizeof( Real ) );
    constexpr int SBLOCK = sblock<Real>();
    constexpr int MAXM = SCATTER_SLOTS;
    constexpr int NPT = RLEN - rhead<Real>();
    constexpr int LMAX = SCATTER_TASK_MAX;
    // Pencil lanes: a task's K*K lanes own the (j, k) rows of its stencil.
    constexpr int KK = K * K;
    constexpr int GPW = 64 / KK;
    constexpr int NGROUP = SBLOCK / 64 * GPW;
    extern __shared__ __align__( 16 ) unsigned char smem[];

    // LDS layout (capacities fixed on the host for a full brick): FP64
    // window [node][DT], base-cell offsets, task table (sorted, unsorted,
    // length slots, length histogram), sorted record order, records, two run
    // tables, part-table entries, scan scratch.
    double* win = reinterpret_cast<double*>( smem );
    size_t off = align16( sizeof( double ) * (size_t)DT * ta.max_win_nodes );
    int* cnt = reinterpret_cast<int*>( smem + off );
    off = align16( off + sizeof( int ) * ( ta.max_base_cells + 1 ) );
    unsigned* task = reinterpret_cast<unsigned*>( smem + off );
    off = align16( off + sizeof( unsigned ) * ta.max_tasks );
    unsigned* task_in = reinterpret_cast<unsigned*>( smem + off );
    off = align16( off + sizeof( unsigned ) * ta.max_tasks );
    int* task_slot = reinterpret_cast<int*>( smem + off );
    off = align16( off + sizeof( int ) * ta.max_tasks );
    int* len_hist = reinterpret_cast<int*>( smem + off );
    int* len_off = len_hist + LMAX;
    off = align16( off + sizeof( int ) * 2 * LMAX );
    uint16_t* ord = reinterpret_cast<uint16_t*>( smem + off );
    off = align16( off + sizeof( uint16_t ) * ta.cap );
    Real* rec = reinterpret_cast<Real*>( smem + off );
    off = align16( off + sizeof( Real ) * (size_t)RS * ta.cap );
    int64_t* run_start[2];
    int* run_pre[2];
    for ( int b = 0; b < 2; ++b )
    {
        run_start[b] = reinterpret_cast<int64_t*>( smem + off );
        off = align16( off + sizeof( int64_t ) * ta.max_runs );
        run_pre[b] = reinterpret_cast<int*>( smem + off );
        off = align16( off + sizeof( int ) * ( ta.max_runs + 1 ) );
    }
    int2* ptab = reinterpret_cast<int2*>( smem + off );
    off = align16( off + sizeof( int2 ) * 128 );
    int* wsum = reinterpret_cast<int*>( smem + off );

    const int W[3] = { ta.win[0], ta.win[1], ta.win[2] };
    const int wnodes = W[0] * W[1] * W[2];
    const float rW[3] = { 1.0f / (float)W[0], 1.0f / (float)W[1], 1.0f / (float)W[2] };
    const PartLoop pl( ta );
    auto part_at = [&]( int l ) {
        const int w = pl.xcd * pl.per_xcd + l;
        return ( l < pl.per_xcd && w < pl.total ) ? w : -1;
    };
    // This workgroup's part-table entries, 64 at a time, in LDS slots
    // k & 127 (k = the workgroup's part counter), read by wave 0.
    auto load_parts = [&]( int k0 ) {
        if ( threadIdx.x < 64 )
        {
            const int k = k0 + (int)threadIdx.x;
            const int wk = part_at( pl.first + k * pl.stride );
            ptab[k & 127] = wk >= 0 ? ta.parts[wk] : make_int2( -1, -1 );
        }
    };
    int loc = pl.first;
    if ( part_at( loc ) < 0 )
        return;

    // The lane's task group and stencil row, and the weight polynomials of
    // its row's y and z nodes.
    const int lane = (int)threadIdx.x & 63;
    const int gw = lane / KK;
    const int lg = lane - gw * KK;
    const int pj = lg / K, pk = lg - ( lg / K ) * K;
    const bool pencil_lane = gw < GPW;
    const int gid = ( (int)threadIdx.x >> 6 ) * GPW + gw;
    const Real rdx0 = (Real)a.mesh.rdx[0];
    NodePoly<Real> py, pz;
    py.init( pj, (Real)a.mesh.rdx[1] );
    pz.init( pk, (Real)a.mesh.rdx[2] );
    const int wstride_i = W[1] * W[2] * DT;

    // Prologue: the first part's run table and particle loads; zeroed
    // window and counts; the first 64 part-table entries.
    load_parts( 0 );
    TileGeom tg;
    part_geometry_of( ta, a.mesh, ta.parts[part_at( loc )], tg );
    if ( threadIdx.x < 64 )
    {
        RunLoad rl;
        runs_issue( ta, a.mesh, tg, rl );
        runs_commit( rl, tg, run_start[0], run_pre[0] );
    }
    for ( int e = threadIdx.x; e < DT * wnodes; e += SBLOCK )
        win[e] = 0.0;
    for ( int c = threadIdx.x; c <= ta.max_base_cells; c += SBLOCK )
        cnt[c] = 0;
    if ( threadIdx.x < LMAX )
        len_hist[threadIdx.x] = 0;
    __syncthreads();
    int buf = 0;
    PartPrefetch<Real, MAXM, NPT> pf;
    {
        const int nruns = ( tg.c1[0] - tg.c0[0] ) * ( tg.c1[1] - tg.c0[1] );
        const int pbeg = tg.part * ta.cap;
        prefetch_part<Real, CONTIG, MAXM, SBLOCK, FORMS...>(
            a, ta, run_start[0], run_pre[0], nruns, pbeg,
            min( ta.cap, run_pre[0][nruns] - pbeg ), pf );
    }
    for ( int kpart = 0;; ++kpart )
    {
        // The next part: its run table is read now and committed after B.
        const int loc1 = loc + pl.stride;
        const bool more = part_at( loc1 ) >= 0;
        TileGeom tg1 = tg;
        RunLoad rl1 = { 0, 0, 0 };
        if ( more )
            part_geometry_of( ta, a.mesh, ptab[( kpart + 1 ) & 127], tg1 );
        // Base cells of the brick and the node window they reach.
        int blo[3], NB[3];
#pragma unroll
        for ( int d = 0; d < 3; ++d )
        {
            blo[d] = tg.c0[d] + a.mesh.halo + BASE_SHIFT;
            NB[d] = ( tg.c1[d] - tg.c0[d] ) + BASE_SPREAD;
        }
        const int nbt = NB[0] * NB[1] * NB[2];

        // A: every particle of the part (loaded into registers during the
        // previous part) gets its stencil base cell, counted per base cell
        // (the slot is the particle's rank within it), and its record (x
        // weights and gradients, y and z cell offsets, premultiplied
        // attributes) is written to LDS at its part index.
        int key[MAXM];
#pragma unroll
        for ( int mi = 0; mi < MAXM; ++mi )
        {
            key[mi] = -1;
            if ( pf.p[mi] < 0 )
                continue;
            const int i = (int)threadIdx.x + mi * SBLOCK;
            // key: -1 no particle or outside the mesh, -2 outside the brick,
            // else base cell << 16 | rank.
            int bs[3];
            Real v[RLEN];
            Real fr[3];
            bool ok = true;
#pragma unroll
            for ( int d = 0; d < 3; ++d )
                ok = spline_frac<ORDER, Real>( logical_coord<ORDER>( pf.x[mi][d], a.mesh.low[d],
                                                                     a.mesh.rdx[d] ),
                                               a.mesh.num_node[d], bs[d], fr[d] ) &&
                     ok;
            if constexpr ( x_weights_in_record<Real>() )
            {
                Spline1D<ORDER, Real> sx;
                spline_weights<ORDER, Real, false>( fr[0], (Real)a.mesh.rdx[0], sx );
#pragma unroll
                for ( int k = 0; k < K; ++k )
                {
                    v[k] = sx.w[k];
                    v[K + k] = sx.g[k];
                }
                v[2 * K] = fr[1];
                v[2 * K + 1] = fr[2];
            }
            else
            {
#pragma unroll
                for ( int d = 0; d < 3; ++d )
                    v[d] = fr[d];
            }
            if ( !ok )
            {
                atomicAdd( a.skipped, 1ull );
                continue;
            }
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
                key[mi] = ( cell << 16 ) | atomicAdd( cnt + cell, 1 );
                static_for<NT>( [&]( auto I ) {
                    constexpr int t = I;
                    constexpr int NP = np_of( FL::v[t] );
                    const Real m = (Real)a.t[t].mult;
#pragma unroll
                    for ( int c = 0; c < NP; ++c )
                        v[RL::attr( t ) + c] = pf.attr[mi][RL::attr( t ) - rhead<Real>() + c] * m;
                } );
                detail::vec_copy_out<Real, RLEN>( v, rec + (size_t)RS * i );
            }
            else
            {
                key[mi] = -2;
            }
        }
        // Particles not sorted into this brick (a binning older than the
        // positions) are added straight to the grid, after the common path
        // so that its registers are free.
#pragma unroll
        for ( int mi = 0; mi < MAXM; ++mi )
            if ( key[mi] == -2 )
            {
                Spline3D<ORDER, Real, false> sp;
                eval_spline( a.mesh, pf.x[mi], sp );
                global_scatter_particle<Real, SHARED, FORMS...>( a, pf.p[mi], sp );
            }
        // The next part's run table reads, committed after B.
        if ( more )
            runs_issue( ta, a.mesh, tg1, rl1 );
        __syncthreads();
        // B: base-cell offsets (cnt[nbt] = sorted particles), then the task
        // table: every base cell's c records cut into ceil(c / LMAX) runs of
        // near-equal length.  One scan of (count | tasks << 16) per base cell
        // gives both the record offsets (low half) and the task offsets (high
        // half); the halves cannot carry into each other (a part holds < 2^15
        // records).  Each task also takes a slot in a histogram of task
        // lengths (longest first), by which it is placed below.
        const int packed_total = block_exclusive_scan<SBLOCK>(
            cnt, nbt, wsum, []( int c ) { return c | ( ( c + LMAX - 1 ) / LMAX ) << 16; } );
        const int ntask = packed_total >> 16;
        for ( int b = threadIdx.x; b < nbt; b += SBLOCK )
        {
            const int e0 = cnt[b];
            const int e1 = b + 1 < nbt ? cnt[b + 1] : packed_total;
            const int c0 = e0 & 0xffff, c = ( e1 & 0xffff ) - c0;
            const int lx = b / ( NB[1] * NB[2] );
            const int ly = ( b / NB[2] ) % NB[1];
            const int lz = b % NB[2];
            int q = e0 >> 16;
            const int nt = ( c + LMAX - 1 ) / LMAX;
            const int len = nt > 0 ? c / nt : 0, longer = nt > 0 ? c - len * nt : 0;
            for ( int i = 0, r = 0; i < nt; ++i )
            {
                const int n = len + ( i < longer ? 1 : 0 );
                const int bucket = LMAX - n;
                task_in[q] = pack_task( lx, ly, lz, c0 + r, n );
                task_slot[q] = ( bucket << 16 ) | atomicAdd( len_hist + bucket, 1 );
                ++q;
                r += n;
            }
        }
        if ( more )
            runs_commit( rl1, tg1, run_start[buf ^ 1], run_pre[buf ^ 1] );
        // C: the sorted order of the records (base cell, then rank).
#pragma unroll
        for ( int mi = 0; mi < MAXM; ++mi )
            if ( key[mi] >= 0 )
                ord[( cnt[key[mi] >> 16] & 0xffff ) + ( key[mi] & 0xffff )] =
                    (uint16_t)( threadIdx.x + mi * SBLOCK );
        __syncthreads();
        // The task table in order of decreasing length: offsets of the length
        // buckets (one wave; the histogram is cleared for the next part),
        // then every task to its bucket's offset plus its slot.
        if ( threadIdx.x < 64 )
        {
            const int lane = threadIdx.x;
            const int h = lane < LMAX ? len_hist[lane] : 0;
            int v = h;
#pragma unroll
            for ( int o = 1; o < 64; o <<= 1 )
            {
                const int u = __shfl_up( v, o, 64 );
                if ( lane >= o )
                    v += u;
            }
            if ( lane < LMAX )
            {
                len_off[lane] = v - h;
                len_hist[lane] = 0;
            }
        }
        __syncthreads();
        for ( int q = threadIdx.x; q < ntask; q += SBLOCK )
        {
            const int sl = task_slot[q];
            task[len_off[sl >> 16] + ( sl & 0xffff )] = task_in[q];
        }
        __syncthreads();
        // D: a task's K*K lanes each own one (j, k) row of the stencil: they
        // walk the task's records (the same record across the group: an
        // LDS broadcast), accumulate the K nodes of their row in registers
        // for every term, and add them once into the FP64 window.
        // Tasks go to the groups in rounds, longest first, alternately
        // forwards and backwards over the groups, so the groups' total
        // lengths stay close.
        if ( pencil_lane )
        {
            for ( int round = 0;; ++round )
            {
                const int q = round * NGROUP + ( ( round & 1 ) ? NGROUP - 1 - gid : gid );
                if ( q >= ntask )
                    break;
                const unsigned e = task[q];
                const int lx = (int)( e >> 28 );
                const int ly = (int)( ( e >> 24 ) & 15 );
                const int lz = (int)( ( e >> 20 ) & 15 );
                const int r0 = (int)( ( e >> 5 ) & 0x7fff );
                const int n = (int)( e & 31 ) + 1;
                Real acc[K][DT];
#pragma unroll
                for ( int i = 0; i < K; ++i )
#pragma unroll
                    for ( int d = 0; d < DT; ++d )
                        acc[i][d] = Real( 0 );
                for ( int r = r0; r < r0 + n; ++r )
                {
                    Real cur[RLEN];
                    read_record<Real, RLEN>( rec + (size_t)RS * ord[r], cur );
                    pencil_record<Real, SHARED, FORMS...>( cur, rdx0, py, pz, acc );
                }
                double* wb = win + ( ( lx * W[1] + ly + pj ) * W[2] + lz + pk ) * DT;
#pragma unroll
                for ( int i = 0; i < K; ++i )
#pragma unroll
                    for ( int d = 0; d < DT; ++d )
                        atomicAdd( wb + i * wstride_i + d, (double)acc[i][d] );
            }
        }
        // The next part's particle reads are issued here, after the
        // accumulation (whose registers they would otherwise share), and
        // complete while this part's window is flushed.
        if ( more )
        {
            const int nruns1 = ( tg1.c1[0] - tg1.c0[0] ) * ( tg1.c1[1] - tg1.c0[1] );
            const int pbeg1 = tg1.part * ta.cap;
            prefetch_part<Real, CONTIG, MAXM, SBLOCK, FORMS...>(
                a, ta, run_start[buf ^ 1], run_pre[buf ^ 1], nruns1, pbeg1,
                min( ta.cap, run_pre[buf ^ 1][nruns1] - pbeg1 ), pf );
        }
        __syncthreads();

        // E: flush the window, one thread per node (the grid's unit-stride
        // dimension fastest across threads): the node's dofs are read and
        // zeroed for the next part, and every non-zero value is added to its
        // grid with one global atomic.
        for ( int f = threadIdx.x; f < wnodes; f += SBLOCK )
        {
            int q[3], d0;
            window_coords<1>( ta, f, W, rW, q, d0 );
            double* src = win + ( ( q[0] * W[1] + q[1] ) * W[2] + q[2] ) * DT;
            double v[DT];
#pragma unroll
            for ( int d = 0; d < DT; ++d )
            {
                v[d] = src[d];
                src[d] = 0.0;
            }
            const int gi = blo[0] + q[0], gj = blo[1] + q[1], gk = blo[2] + q[2];
            if ( gi < 0 || gj < 0 || gk < 0 || gi >= a.mesh.num_node[0] ||
                 gj >= a.mesh.num_node[1] || gk >= a.mesh.num_node[2] )
                continue;
            static_for<( SHARED ? 1 : NT )>( [&]( auto I ) {
                constexpr int t = I;
                constexpr int NG = SHARED ? 3 : ng_of( FL::v[t] );
                constexpr int O = SD::offset( t, SHARED );
                const GFieldDev& g = a.t[t].grid;
                Real* dst = gptr<Real>( g, gi * g.s[0] + gj * g.s[1] + gk * g.s[2] );
#pragma unroll
                for ( int d = 0; d < NG; ++d )
                {
                    const Real x = (Real)v[O + d];
                    if ( x != Real( 0 ) )
                        atomic_add( dst + d * g.s[3], x );
                }
            } );
        }
        for ( int c = threadIdx.x; c <= ta.max_base_cells; c += SBLOCK )
            cnt[c] = 0;
        if ( ( ( kpart + 2 ) & 63 ) == 0 )
            load_parts( kpart + 2 );
        // The next part reuses the window, the counts and the tables.
        __syncthreads();
        if ( !more )
            break;
        loc = loc1;
        tg = tg1;
        buf ^= 1;
    }
}

// ---------------------------------------------------------------------------
// Launch helpers
// ---------------------------------------------------------------------------

template <class KernelPtr, class Args>
hipError_t launch_tile_kernel( KernelPtr k, const Args& a, const TileArgs& t,
                               size_t lds, hipStream_t s, int block = BLOCK )
{
    if ( lds > 64 * 1024 )
    {
        hipError_t e = hipFuncSetAttribute(
            reinterpret_cast<const void*>( k ),
            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds );
        if ( e != hipSuccess )
            return e;
    }
    hipLaunchKernelGGL( k, dim3( (unsigned)t.grid ), dim3( block ), lds, s, a, t );
    return hipGetLastError();
}

template <class Real, bool CONTIG>
hipError_t gather_tiled_launch( int plan, int mode, const GatherArgs& a, const TileArgs& t,
                                size_t lds, hipStream_t s )
{
    if ( plan == PLAN_SINGLE )
    {
        switch ( a.t[0].form )
        {
#define CA_CASE( F )                                                           \
    case F:                                                                    \
        return mode == GATHER_OUT                                              \
                   ? launch_tile_kernel( gather_tiled_kernel<Real, GATHER_OUT, CONTIG, false, F>, \
                                         a, t, lds, s )                        \
               : mode == GATHER_OUT_XBAR                                       \
                   ? launch_tile_kernel(                                       \
                         gather_tiled_kernel<Real, GATHER_OUT_XBAR, CONTIG, false, F>, a, t, \
                         lds, s )                                              \
                   : launch_tile_kernel( gather_tiled_kernel<Real, GATHER_XBAR, CONTIG, false, F>, \
                                         a, t, lds, s );
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
    }
    if ( plan == PLAN_MPM_G2P && mode == GATHER_OUT )
        return launch_tile_kernel(
            gather_tiled_kernel<Real, GATHER_OUT, CONTIG, true, FORM_VV, FORM_TDT>, a, t, lds, s );
    if ( plan == PLAN_MPM_G2P && mode == GATHER_XBAR )
        return launch_tile_kernel(
            gather_tiled_kernel<Real, GATHER_XBAR, CONTIG, true, FORM_VV, FORM_TDT>, a, t, lds, s );
    if ( plan == PLAN_MPM_P2G && mode == GATHER_OUT_XBAR )
        return launch_tile_kernel(
            gather_tiled_kernel<Real, GATHER_OUT_XBAR, CONTIG, false, FORM_SV, FORM_VV, FORM_TD>,
            a, t, lds, s );
    if ( plan == PLAN_MPM_ADJOINT && mode == GATHER_OUT_XBAR )
        return launch_tile_kernel(
            gather_tiled_kernel<Real, GATHER_OUT_XBAR, CONTIG, false, FORM_SV, FORM_VV, FORM_TD,
                                FORM_VV | FORM_XONLY,
                                FORM_TDT | FORM_XONLY | FORM_SHARE_PREV>,
            a, t, lds, s );
    return hipErrorInvalidValue;
}

template <class Real>
hipError_t gather_tiled_typed( Launcher& L, int plan, const GatherArgs& a,
                               const TileArgs& t, size_t lds, hipStream_t s )
{
    const int mode = !a.has_xbar ? GATHER_OUT : ( a.t[0].pout.data ? GATHER_OUT_XBAR : GATHER_XBAR );
    const char* name =
        plan == PLAN_MPM_ADJOINT
            ? "adjoint_mpm_tiled"
            : a.has_xbar ? ( plan == PLAN_MPM_P2G ? "p2g_adjoint_mpm_tiled" : "gather_xbar_tiled" )
                         : ( plan == PLAN_MPM_G2P ? "g2p_mpm_tiled" : "g2p_tiled" );
    L.before( name, s );
    const hipError_t e = a.contig ? gather_tiled_launch<Real, true>( plan, mode, a, t, lds, s )
                                  : gather_tiled_launch<Real, false>( plan, mode, a, t, lds, s );
    L.after( name, s );
    return e;
}

template <class Real, bool CONTIG>
hipError_t scatter_cells_launch( int plan, const ScatterArgs& a, const TileArgs& t, size_t lds,
                                 hipStream_t s )
{
    if ( plan == PLAN_SINGLE )
    {
        switch ( a.t[0].form )
        {
#define CA_CASE( F )                                                           \
    case F:                                                                    \
        return launch_tile_kernel( scatter_cells_kernel<Real, CONTIG, false, F>, a, t, lds, s, sblock<Real>() );
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
    }
    if ( plan == PLAN_MPM_P2G )
        return launch_tile_kernel(
            scatter_cells_kernel<Real, CONTIG, false, FORM_SV, FORM_VV, FORM_TD>, a, t, lds, s,
            sblock<Real>() );
    if ( plan == PLAN_MASS_MOMENTUM )
        return launch_tile_kernel( scatter_cells_kernel<Real, CONTIG, false, FORM_SV, FORM_VV>,
                                   a, t, lds, s, sblock<Real>() );
    if ( plan == PLAN_MPM_G2P )
        return launch_tile_kernel( scatter_cells_kernel<Real, CONTIG, true, FORM_VV, FORM_TDT>,
                                   a, t, lds, s, sblock<Real>() );
    return hipErrorInvalidValue;
}

template <class Real>
hipError_t scatter_cells_typed( Launcher& L, int plan, const ScatterArgs& a,
                                const TileArgs& t, size_t lds, hipStream_t s )
{
    const char* name = plan == PLAN_MPM_P2G         ? "p2g_mpm_cells"
                       : plan == PLAN_MPM_G2P       ? "g2p_adjoint_mpm_cells"
                       : plan == PLAN_MASS_MOMENTUM ? "p2g_mass_momentum_cells"
                                                    : "scatter_cells";
    L.before( name, s );
    const hipError_t e = a.contig ? scatter_cells_launch<Real, true>( plan, a, t, lds, s )
                                  : scatter_cells_launch<Real, false>( plan, a, t, lds, s );
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
