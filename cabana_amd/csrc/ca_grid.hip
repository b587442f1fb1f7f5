// Grid-side kernels: the single-rank periodic node halo (fill and fold) and
// the cell binning (LinkedCellList build + permute) used to order particles.
#include "ca_internal.hpp"

namespace cabana_amd
{
namespace
{

constexpr int BLOCK = 256;

inline dim3 blocks_for( int64_t n )
{
    return dim3( (unsigned)( ( n + BLOCK - 1 ) / BLOCK ) );
}

// Owned entity range of one dimension on one rank
// (Cabana_Grid_LocalGrid_impl.hpp:239-280, 430-450, 534-560): N entities,
// plus the closing boundary node in a non-periodic node-centred dimension.
__device__ __forceinline__ int owned_hi( const MeshDev& m, int d )
{
    return m.halo + m.num_cell[d] + ( ( m.periodic[d] || m.stag[d] ) ? 0 : 1 );
}

// Classifies node index x of dimension d for a halo of width w.  Returns the
// owner index (itself when owned), or -1 when the node is a ghost that no
// message of the 26-direction exchange reaches.  The bands are the shared
// index spaces of Cabana_Grid_LocalGrid_impl.hpp:370-398, 470-509: low
// ghosts [h - w, h) receive from [h + N - w, h + N), high ghosts
// [h + N, h + N + w + p) receive from [h, h + w + p), p = 1 in a
// node-centred dimension and 0 in a cell-centred one.
__device__ __forceinline__ int owner_of( const MeshDev& m, int d, int x, int w,
                                         bool& is_ghost )
{
    const int h = m.halo, N = m.num_cell[d];
    is_ghost = false;
    if ( x >= h && x < owned_hi( m, d ) )
        return x;
    is_ghost = true;
    if ( !m.periodic[d] )
        return -1;
    if ( x >= h - w && x < h )
        return x + N;
    if ( x >= h + N && x < h + N + w + 1 - m.stag[d] )
        return x - N;
    return -1;
}

template <class Real>
__global__ __launch_bounds__( BLOCK ) void halo_fill_kernel( MeshDev m,
                                                             GFieldDev g,
                                                             int w,
                                                             int zero_other )
{
    const int64_t n = (int64_t)g.ext[0] * g.ext[1] * g.ext[2];
    const int64_t e = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if ( e >= n )
        return;
    const int i = (int)( e / ( (int64_t)g.ext[1] * g.ext[2] ) );
    const int j = (int)( ( e / g.ext[2] ) % g.ext[1] );
    const int k = (int)( e % g.ext[2] );
    bool gi, gj, gk;
    const int oi = owner_of( m, 0, i, w, gi );
    const int oj = owner_of( m, 1, j, w, gj );
    const int ok = owner_of( m, 2, k, w, gk );
    if ( !gi && !gj && !gk )
        return; // owned
    Real* dst = gptr<Real>( g, i * g.s[0] + j * g.s[1] +
                k * g.s[2] );
    if ( oi < 0 || oj < 0 || ok < 0 )
    {
        if ( zero_other )
            for ( int d = 0; d < g.ext[3]; ++d )
                dst[d * g.s[3]] = Real( 0 );
        return;
    }
    const Real* src = gptr<const Real>( g, oi * g.s[0] +
                      oj * g.s[1] + ok * g.s[2] );
    for ( int d = 0; d < g.ext[3]; ++d )
        dst[d * g.s[3]] = src[d * g.s[3]];
}

// Fold: every owned node adds all of its periodic ghost images.  Each owner
// gathers its images itself, so the sum needs no atomics and its order is
// fixed (the reference's MPI Waitany order is not).
template <class Real>
__global__ __launch_bounds__( BLOCK ) void halo_fold_kernel( MeshDev m,
                                                             GFieldDev g,
                                                             int w )
{
    int lo[3], n[3];
#pragma unroll
    for ( int d = 0; d < 3; ++d )
    {
        lo[d] = m.halo;
        n[d] = owned_hi( m, d ) - m.halo;
    }
    const int64_t total = (int64_t)n[0] * n[1] * n[2];
    const int64_t e = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if ( e >= total )
        return;
    const int x[3] = { lo[0] + (int)( e / ( (int64_t)n[1] * n[2] ) ),
                       lo[1] + (int)( ( e / n[2] ) % n[1] ),
                       lo[2] + (int)( e % n[2] ) };
    int cand[3][3], nc[3];
    bool any = false;
#pragma unroll
    for ( int d = 0; d < 3; ++d )
    {
        const int h = m.halo, N = m.num_cell[d];
        cand[d][0] = x[d];
        nc[d] = 1;
        if ( m.periodic[d] )
        {
            if ( x[d] < h + w + 1 - m.stag[d] )
                cand[d][nc[d]++] = x[d] + N;
            if ( x[d] >= h + N - w )
                cand[d][nc[d]++] = x[d] - N;
        }
        any = any || nc[d] > 1;
    }
    if ( !any )
        return;
    Real* dst = gptr<Real>( g, x[0] * g.s[0] +
                x[1] * g.s[1] + x[2] * g.s[2] );
    for ( int d = 0; d < g.ext[3]; ++d )
    {
        Real acc = Real( 0 );
        for ( int a = 0; a < nc[0]; ++a )
            for ( int b = 0; b < nc[1]; ++b )
                for ( int c = 0; c < nc[2]; ++c )
                {
                    if ( a == 0 && b == 0 && c == 0 )
                        continue;
                    acc += *gptr<const Real>(
                        g, cand[0][a] * g.s[0] + cand[1][b] * g.s[1] +
                               cand[2][c] * g.s[2] + d * g.s[3] );
                }
        dst[d * g.s[3]] += acc;
    }
}

template <class Real>
__global__ __launch_bounds__( BLOCK ) void zero_ghosts_kernel( MeshDev m,
                                                               GFieldDev g )
{
    const int64_t n = (int64_t)g.ext[0] * g.ext[1] * g.ext[2];
    const int64_t e = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if ( e >= n )
        return;
    const int x[3] = { (int)( e / ( (int64_t)g.ext[1] * g.ext[2] ) ),
                       (int)( ( e / g.ext[2] ) % g.ext[1] ),
                       (int)( e % g.ext[2] ) };
    bool owned = true;
#pragma unroll
    for ( int d = 0; d < 3; ++d )
        owned = owned && x[d] >= m.halo && x[d] < owned_hi( m, d );
    if ( owned )
        return;
    Real* dst = gptr<Real>( g, x[0] * g.s[0] +
                x[1] * g.s[1] + x[2] * g.s[2] );
    for ( int d = 0; d < g.ext[3]; ++d )
        dst[d * g.s[3]] = Real( 0 );
}

// ---------------------------------------------------------------------------
// Binning
// ---------------------------------------------------------------------------

// locatePoint + cardinalCellIndex (Cabana_CartesianGrid.hpp:171-182,
// 226-231, 279-283) with the same FP64 rounding: floor((x - min) * rdx),
// clamped only when it equals nx.
__device__ __forceinline__ int64_t locate_cell( const CartGridDev& g,
                                                const double p[3] )
{
#pragma clang fp contract( off )
    int c[3];
#pragma unroll
    for ( int d = 0; d < 3; ++d )
    {
        const double f = floor( ( p[d] - g.min[d] ) * g.rdx[d] );
        if ( !( f >= 0.0 && f <= (double)g.nx[d] ) )
            return -1;
        int v = (int)f;
        v = ( v == g.nx[d] ) ? v - 1 : v;
        c[d] = v;
    }
    return ( (int64_t)c[0] * g.nx[1] + c[1] ) * g.nx[2] + c[2];
}

// Bin key of every particle in [begin, end): its cardinal cell, or bad_key
// (one past the last cell) when the particle lies outside the grid.  No
// per-cell counters are touched here: the counts come out of the sort.
__global__ __launch_bounds__( BLOCK ) void lcl_key_kernel(
    PFieldDev pos, int64_t begin, int64_t end, CartGridDev g, int32_t bad_key,
    int32_t* keys, unsigned long long* bad )
{
    const int64_t p = begin + (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if ( p >= end )
        return;
    const uint32_t b = pos.base( p );
    const double x[3] = { *pos.ptr<const double>( b, 0 ),
                          *pos.ptr<const double>( b, 1 ),
                          *pos.ptr<const double>( b, 2 ) };
    const int64_t c = locate_cell( g, x );
    keys[p - begin] = c < 0 ? bad_key : (int32_t)c;
    const unsigned long long out = __ballot( c < 0 );
    if ( out && (int)( threadIdx.x & 63 ) == __ffsll( (long long)out ) - 1 )
        atomicAdd( bad, (unsigned long long)__popcll( out ) );
}

constexpr int SCAN_BLOCK = 1024;

__global__ __launch_bounds__( SCAN_BLOCK ) void scan_reduce_kernel(
    const int32_t* in, int64_t n, int64_t* block_sums )
{
    __shared__ int64_t red[SCAN_BLOCK / 64];
    const int64_t i = (int64_t)blockIdx.x * SCAN_BLOCK + threadIdx.x;
    int64_t v = ( i < n ) ? in[i] : 0;
    for ( int o = 32; o > 0; o >>= 1 )
        v += __shfl_down( v, o, 64 );
    if ( ( threadIdx.x & 63 ) == 0 )
        red[threadIdx.x >> 6] = v;
    __syncthreads();
    if ( threadIdx.x == 0 )
    {
        int64_t s = 0;
        for ( int w = 0; w < SCAN_BLOCK / 64; ++w )
            s += red[w];
        block_sums[blockIdx.x] = s;
    }
}

// Exclusive scan of the block sums by one workgroup (sequential per thread
// chunk, then a workgroup scan of the chunk totals).
__global__ __launch_bounds__( SCAN_BLOCK ) void scan_block_sums_kernel(
    int64_t* block_sums, int64_t nb )
{
    __shared__ int64_t part[SCAN_BLOCK];
    const int64_t chunk = ( nb + SCAN_BLOCK - 1 ) / SCAN_BLOCK;
    const int64_t lo = threadIdx.x * chunk;
    const int64_t hi = lo + chunk < nb ? lo + chunk : nb;
    int64_t s = 0;
    for ( int64_t i = lo; i < hi; ++i )
        s += block_sums[i];
    part[threadIdx.x] = s;
    __syncthreads();
    if ( threadIdx.x == 0 )
    {
        int64_t run = 0;
        for ( int t = 0; t < SCAN_BLOCK; ++t )
        {
            const int64_t v = part[t];
            part[t] = run;
            run += v;
        }
    }
    __syncthreads();
    int64_t run = part[threadIdx.x];
    for ( int64_t i = lo; i < hi; ++i )
    {
        const int64_t v = block_sums[i];
        block_sums[i] = run;
        run += v;
    }
}

__global__ __launch_bounds__( SCAN_BLOCK ) void scan_apply_kernel(
    const int32_t* in, int64_t* out, int64_t n, const int64_t* block_sums )
{
    __shared__ int64_t wsum[SCAN_BLOCK / 64];
    const int64_t i = (int64_t)blockIdx.x * SCAN_BLOCK + threadIdx.x;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int64_t x = ( i < n ) ? in[i] : 0;
    // Inclusive wave scan.
    int64_t v = x;
    for ( int o = 1; o < 64; o <<= 1 )
    {
        const int64_t u = __shfl_up( v, o, 64 );
        if ( lane >= o )
            v += u;
    }
    if ( lane == 63 )
        wsum[wv] = v;
    __syncthreads();
    if ( threadIdx.x == 0 )
    {
        int64_t run = 0;
        for ( int w = 0; w < SCAN_BLOCK / 64; ++w )
        {
            const int64_t t = wsum[w];
            wsum[w] = run;
            run += t;
        }
    }
    __syncthreads();
    if ( i < n )
        out[i] = block_sums[blockIdx.x] + wsum[wv] + v - x;
}

// ---------------------------------------------------------------------------
// Stable LSD radix sort of the bin keys.  Cabana's Serial build leaves the
// particles of one bin in ascending index order
// (Cabana_LinkedCellList.hpp:710-731, one thread); a stable sort by key gives
// that order directly, with no per-bin atomics and no per-bin sorting, so
// crowded bins cost the same as sparse ones.
//
// A tile of RS_TILE keys belongs to one workgroup; wave w owns the contiguous
// run [w * 64 * RS_ITEMS, (w + 1) * 64 * RS_ITEMS) of its tile and walks it
// 64 keys at a time, so (wave, step, lane) is index order.  Pass layout:
//   hist:    per-tile digit counts, digit-major: hist[digit * ntiles + tile];
//   scan:    exclusive scan of hist gives every (digit, tile) its output base;
//   scatter: each key's rank among equal digits of its tile (ballot peers +
//            per-wave running counts in LDS), a tile-local reorder in LDS,
//            then writes in tile order, so equal digits leave contiguously.
// ---------------------------------------------------------------------------
constexpr int RS_BLOCK = 256;
constexpr int RS_ITEMS = 16;
constexpr int RS_TILE = RS_BLOCK * RS_ITEMS;
constexpr int RS_WAVES = RS_BLOCK / 64;
constexpr int RS_MAX_BITS = 8;
constexpr int RS_MAX_BINS = 1 << RS_MAX_BITS;
static_assert( RS_MAX_BINS == RS_BLOCK, "one digit per thread in the scans" );

__global__ __launch_bounds__( RS_BLOCK ) void radix_hist_kernel(
    const int32_t* keys, int64_t n, int shift, int nbins, int64_t ntiles,
    int32_t* hist )
{
    __shared__ uint32_t h[RS_MAX_BINS];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int64_t t0 = (int64_t)blockIdx.x * RS_TILE;
    const uint32_t mask = (uint32_t)nbins - 1u;
#pragma unroll 4
    for ( int k = 0; k < RS_ITEMS; ++k )
    {
        const int64_t i = t0 + k * RS_BLOCK + threadIdx.x;
        if ( i < n )
            atomicAdd( &h[( (uint32_t)keys[i] >> shift ) & mask], 1u );
    }
    __syncthreads();
    if ( (int)threadIdx.x < nbins )
        hist[threadIdx.x * ntiles + blockIdx.x] = (int32_t)h[threadIdx.x];
}

// Exclusive scan of one value per thread over the workgroup; returns the
// prefix of this thread and leaves the total in *total.
__device__ __forceinline__ uint32_t rs_block_scan( uint32_t v, uint32_t* wsum,
                                                   uint32_t* total )
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for ( int o = 1; o < 64; o <<= 1 )
    {
        const uint32_t u = __shfl_up( x, o, 64 );
        if ( lane >= o )
            x += u;
    }
    if ( lane == 63 )
        wsum[w] = x;
    __syncthreads();
    uint32_t before = 0, all = 0;
#pragma unroll
    for ( int q = 0; q < RS_WAVES; ++q )
    {
        const uint32_t s = wsum[q];
        before += q < w ? s : 0u;
        all += s;
    }
    *total = all;
    return before + x - v;
}

// One stable pass over bits [shift, shift + dbits).  FIRST: the values are
// the key indices themselves (not read).  LAST: the values leave as
// begin + index into the 64-bit permutation.
template <bool FIRST, bool LAST>
__global__ __launch_bounds__( RS_BLOCK ) void radix_scatter_kernel(
    const int32_t* keys_in, const int32_t* vals_in, int64_t n, int shift,
    int nbins, int64_t ntiles, const int64_t* digit_base, int32_t* keys_out,
    int32_t* vals_out, int64_t* perm_out, int64_t begin )
{
    __shared__ uint32_t wcnt[RS_WAVES][RS_MAX_BINS];
    __shared__ uint32_t tstart[RS_MAX_BINS];
    __shared__ int64_t gadj[RS_MAX_BINS];
    __shared__ uint32_t wsum[RS_WAVES];
    __shared__ int32_t lkey[RS_TILE];
    __shared__ int32_t lval[RS_TILE];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for ( int q = 0; q < RS_WAVES; ++q )
        wcnt[q][threadIdx.x] = 0;
    __syncthreads();
    const int64_t t0 = (int64_t)blockIdx.x * RS_TILE;
    const int64_t wbase = t0 + (int64_t)w * 64 * RS_ITEMS;
    const uint32_t mask = (uint32_t)nbins - 1u;
    const uint64_t below = ( 1ull << lane ) - 1ull;
    int32_t key[RS_ITEMS], val[RS_ITEMS];
    uint32_t rank[RS_ITEMS];
#pragma unroll
    for ( int k = 0; k < RS_ITEMS; ++k )
    {
        const int64_t i = wbase + k * 64 + lane;
        const bool ok = i < n;
        key[k] = ok ? keys_in[i] : 0;
        if constexpr ( FIRST )
            val[k] = (int32_t)i;
        else
            val[k] = ok ? vals_in[i] : 0;
    }
#pragma unroll
    for ( int k = 0; k < RS_ITEMS; ++k )
    {
        const bool ok = wbase + k * 64 + lane < n;
        const uint32_t d = ( (uint32_t)key[k] >> shift ) & mask;
        // Lanes holding the same digit: intersect the ballots of every bit.
        uint64_t peers = __ballot( ok );
#pragma unroll
        for ( int b = 0; b < RS_MAX_BITS; ++b )
        {
            const bool set = ( d >> b ) & 1u;
            const uint64_t bb = __ballot( set );
            peers &= set ? bb : ~bb;
        }
        if ( ok )
        {
            const uint32_t c = wcnt[w][d];
            rank[k] = c + (uint32_t)__popcll( peers & below );
            // The highest lane of the group advances the running count; it
            // writes a value computed from its own read, so the write
            // follows every lane's read of this step.
            if ( ( peers >> lane ) == 1ull )
                wcnt[w][d] = c + (uint32_t)__popcll( peers );
        }
    }
    __syncthreads();
    // Per digit: prefix over the waves, then the tile's digit starts.
    {
        const int d = threadIdx.x;
        uint32_t run = 0;
#pragma unroll
        for ( int q = 0; q < RS_WAVES; ++q )
        {
            const uint32_t c = wcnt[q][d];
            wcnt[q][d] = run;
            run += c;
        }
        uint32_t total;
        const uint32_t st = rs_block_scan( d < nbins ? run : 0u, wsum, &total );
        tstart[d] = st;
        gadj[d] = d < nbins ? digit_base[(int64_t)d * ntiles + blockIdx.x] - st
                            : 0;
    }
    __syncthreads();
#pragma unroll
    for ( int k = 0; k < RS_ITEMS; ++k )
    {
        if ( wbase + k * 64 + lane < n )
        {
            const uint32_t d = ( (uint32_t)key[k] >> shift ) & mask;
            const uint32_t lp = tstart[d] + wcnt[w][d] + rank[k];
            lkey[lp] = key[k];
            lval[lp] = val[k];
        }
    }
    __syncthreads();
    const int cnt = (int)( n - t0 < RS_TILE ? n - t0 : RS_TILE );
    for ( int j = threadIdx.x; j < cnt; j += RS_BLOCK )
    {
        const int32_t k = lkey[j];
        const int64_t o = gadj[( (uint32_t)k >> shift ) & mask] + j;
        keys_out[o] = k;
        if constexpr ( LAST )
            perm_out[o] = begin + lval[j];
        else
            vals_out[o] = lval[j];
    }
}

// counts and offsets of every bin from the sorted keys: offsets[c] is the
// first position whose key is >= c (the exclusive scan of the counts,
// Cabana_LinkedCellList.hpp:700-705), counts[c] the length of the run.
__device__ __forceinline__ int64_t lower_bound_i32( const int32_t* a,
                                                    int64_t n, int32_t v )
{
    int64_t lo = 0, len = n;
    while ( len > 0 )
    {
        const int64_t half = len >> 1;
        if ( a[lo + half] < v )
        {
            lo += half + 1;
            len -= half + 1;
        }
        else
            len = half;
    }
    return lo;
}

__global__ __launch_bounds__( BLOCK ) void lcl_bounds_kernel(
    const int32_t* sorted_keys, int64_t n, int64_t ncell, int32_t* counts,
    int64_t* offsets )
{
    const int64_t c = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if ( c >= ncell )
        return;
    const int64_t lo = lower_bound_i32( sorted_keys, n, (int32_t)c );
    const int64_t hi = lower_bound_i32( sorted_keys, n, (int32_t)( c + 1 ) );
    offsets[c] = lo;
    counts[c] = (int32_t)( hi - lo );
}

template <class T>
__global__ __launch_bounds__( BLOCK ) void permute_gather_kernel(
    PFieldDev f, const int64_t* permutes, int64_t begin, int64_t end,
    T* scratch )
{
    const int64_t i = begin + (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if ( i >= end )
        return;
    const int64_t src = permutes[i - begin];
    const uint32_t b = f.base( src );
    for ( int c = 0; c < f.ncomp; ++c )
        scratch[( i - begin ) * f.ncomp + c] = *f.ptr<const T>( b, c );
}

template <class T>
__global__ __launch_bounds__( BLOCK ) void permute_copy_back_kernel(
    PFieldDev f, int64_t begin, int64_t end, const T* scratch )
{
    const int64_t i = begin + (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if ( i >= end )
        return;
    const uint32_t b = f.base( i );
    for ( int c = 0; c < f.ncomp; ++c )
        *f.ptr<T>( b, c ) = scratch[( i - begin ) * f.ncomp + c];
}

// Row-contiguous fields (each particle's components in one run of NW 4-byte
// words, rows row_words apart): one thread per word of the permuted range,
// so consecutive threads write consecutive scratch words and read
// consecutive words of each source row.
// Words per thread of the word-parallel permute kernels: a workgroup moves
// PERMUTE_WPT * BLOCK consecutive words, thread t the words t + k * BLOCK,
// so every access of the wave stays coalesced while each thread keeps
// 2 * PERMUTE_WPT independent loads in flight.
constexpr int PERMUTE_WPT = 8;

template <int NW>
__global__ __launch_bounds__( BLOCK ) void permute_gather_words_kernel(
    const uint32_t* src, int64_t row_words, int nw_rt, const int64_t* permutes,
    int64_t nwords, uint32_t* scratch )
{
    const int nw = NW > 0 ? NW : nw_rt;
    // 32-bit index math (the host keeps nwords below 2^31).
    const uint32_t w0 = (uint32_t)blockIdx.x * ( PERMUTE_WPT * BLOCK ) + threadIdx.x;
    int64_t srcw[PERMUTE_WPT];
#pragma unroll
    for ( int k = 0; k < PERMUTE_WPT; ++k )
    {
        const uint32_t w = w0 + k * BLOCK;
        const uint32_t i = w / (uint32_t)nw;
        const int j = (int)( w - i * (uint32_t)nw );
        srcw[k] = w < nwords ? permutes[i] * row_words + j : -1;
    }
    uint32_t v[PERMUTE_WPT];
#pragma unroll
    for ( int k = 0; k < PERMUTE_WPT; ++k )
        v[k] = srcw[k] >= 0 ? src[srcw[k]] : 0u;
#pragma unroll
    for ( int k = 0; k < PERMUTE_WPT; ++k )
        if ( srcw[k] >= 0 )
            scratch[w0 + k * BLOCK] = v[k];
}

// All row-contiguous members of a particle set in one pass: a workgroup
// takes PERMUTE_CHUNK consecutive destination particles, reads their
// permutation entries once into LDS, then copies each member's words of the
// chunk into that member's block of the scratch (member-major: member f at
// scratch + n * word_offset[f]), word-parallel and coalesced as in
// permute_gather_words_kernel.
constexpr int PERMUTE_CHUNK = 512;

__global__ __launch_bounds__( BLOCK ) void permute_gather_members_kernel(
    PermuteMembers pm, const int64_t* permutes, int64_t n, uint32_t* scratch )
{
    __shared__ int64_t perm[PERMUTE_CHUNK];
    const int64_t i0 = (int64_t)blockIdx.x * PERMUTE_CHUNK;
    const int np = (int)min( (int64_t)PERMUTE_CHUNK, n - i0 );
    for ( int i = threadIdx.x; i < np; i += BLOCK )
        perm[i] = permutes[i0 + i];
    __syncthreads();
    for ( int f = 0; f < pm.count; ++f )
    {
        const uint32_t* src = pm.src[f];
        const int64_t rw = pm.row_words[f];
        const uint32_t nw = (uint32_t)pm.words[f];
        uint32_t* dst = scratch + n * pm.word_offset[f] + i0 * nw;
        const uint32_t nwords = (uint32_t)np * nw;
        for ( uint32_t w0 = threadIdx.x; w0 < nwords; w0 += PERMUTE_WPT * BLOCK )
        {
            uint32_t v[PERMUTE_WPT];
#pragma unroll
            for ( int k = 0; k < PERMUTE_WPT; ++k )
            {
                const uint32_t w = w0 + k * BLOCK;
                const uint32_t i = w / nw;
                v[k] = w < nwords ? src[perm[i] * rw + ( w - i * nw )] : 0u;
            }
#pragma unroll
            for ( int k = 0; k < PERMUTE_WPT; ++k )
                if ( w0 + k * BLOCK < nwords )
                    dst[w0 + k * BLOCK] = v[k];
        }
    }
}

// Copy back into padded rows (row_words > words per particle).
template <int NW>
__global__ __launch_bounds__( BLOCK ) void permute_put_words_kernel(
    uint32_t* dst, int64_t row_words, int nw_rt, int64_t begin, int64_t nwords,
    const uint32_t* scratch )
{
    const int nw = NW > 0 ? NW : nw_rt;
    const int64_t w = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if ( w >= nwords )
        return;
    const uint32_t w32 = (uint32_t)w;
    const uint32_t i = w32 / (uint32_t)nw;
    const int j = (int)( w32 - i * (uint32_t)nw );
    dst[( begin + (int64_t)i ) * row_words + j] = scratch[w];
}

// Part table of a brick tiling.  One thread per brick counts the brick's
// particles through the cell offsets of its cell runs and the number of
// parts of at most t.cap particles (an empty brick has none).
__global__ __launch_bounds__( BLOCK ) void brick_count_kernel( TileArgs t, int3 nc,
                                                               int32_t* nparts )
{
    const int nbricks = t.ntile[0] * t.ntile[1] * t.ntile[2];
    const int b = blockIdx.x * BLOCK + threadIdx.x;
    if ( b >= nbricks )
        return;
    const int bi = b / ( t.ntile[1] * t.ntile[2] );
    const int bj = ( b / t.ntile[2] ) % t.ntile[1];
    const int bk = b % t.ntile[2];
    const int i0 = bi * t.tile[0], i1 = min( i0 + t.tile[0], nc.x );
    const int j0 = bj * t.tile[1], j1 = min( j0 + t.tile[1], nc.y );
    const int k0 = bk * t.tile[2], k1 = min( k0 + t.tile[2], nc.z );
    int64_t count = 0;
    for ( int i = i0; i < i1; ++i )
        for ( int j = j0; j < j1; ++j )
        {
            const int64_t first = ( (int64_t)i * nc.y + j ) * nc.z + k0;
            const int64_t last = first + ( k1 - k0 ) - 1;
            count += t.cell_offsets[last] + t.cell_counts[last] - t.cell_offsets[first];
        }
    nparts[b] = (int32_t)( ( count + t.cap - 1 ) / t.cap );
}

__global__ __launch_bounds__( BLOCK ) void part_fill_kernel( const int32_t* nparts,
                                                             const int64_t* start,
                                                             int nbricks, int2* parts,
                                                             int32_t* num_parts )
{
    const int b = blockIdx.x * BLOCK + threadIdx.x;
    if ( b >= nbricks )
        return;
    const int64_t s0 = start[b];
    const int np = nparts[b];
    for ( int q = 0; q < np; ++q )
        parts[s0 + q] = make_int2( b, q );
    if ( b == nbricks - 1 )
        *num_parts = (int32_t)( s0 + np );
}

__global__ __launch_bounds__( BLOCK ) void fill_i32_kernel( int32_t* dst,
                                                            int64_t n,
                                                            int32_t value )
{
    const int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if ( i < n )
        dst[i] = value;
}


// ---------------------------------------------------------------------------
// Particle migration on one rank (Cabana_Grid_ParticleDistributor.hpp)
// ---------------------------------------------------------------------------

// Reports a wave's flagged lanes to one 64-bit counter with one atomic.
__device__ __forceinline__ void count_flagged( bool flag, unsigned long long* counter )
{
    const unsigned long long m = __ballot( flag );
    if ( m && (int)( threadIdx.x & 63 ) == __ffsll( (long long)m ) - 1 )
        atomicAdd( counter, (unsigned long long)__popcll( m ) );
}

// migrateCount (:170-221): particles with a coordinate below the band's low
// or above its high corner (a NaN coordinate compares false, as there).
__global__ __launch_bounds__( BLOCK ) void migrate_count_kernel( PFieldDev pos, int64_t n,
                                                                 MigrateBoundsDev b,
                                                                 unsigned long long* count )
{
    const int64_t p = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if ( p >= n )
        return;
    const uint32_t a = pos.base( p );
    bool out = false;
#pragma unroll
    for ( int d = 0; d < 3; ++d )
    {
        const double x = *pos.ptr<const double>( a, d );
        out = out || x < b.band_lo[d] || x > b.band_hi[d];
    }
    count_flagged( out, count );
}

// getMigrateDestinations (:87-148) on one rank.  Every neighbour of a
// periodic dimension is this rank; past the owned domain of a non-periodic
// dimension there is no neighbour (rank -1), and Distributor removes the
// particle.  The destination is taken from the unshifted position; then each
// periodic coordinate above the global high corner moves down by the extent
// and one below the low corner moves up (one shift, as there).
__global__ __launch_bounds__( BLOCK ) void migrate_shift_kernel( PFieldDev pos, int64_t n,
                                                                 MigrateBoundsDev b,
                                                                 int32_t* keep,
                                                                 unsigned long long* dropped )
{
    const int64_t p = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if ( p >= n )
        return;
    const uint32_t a = pos.base( p );
    bool drop = false;
#pragma unroll
    for ( int d = 0; d < 3; ++d )
    {
        double* xp = pos.ptr<double>( a, d );
        const double x = *xp;
        const bool outside = x < b.own_lo[d] || x > b.own_hi[d];
        if ( b.periodic[d] )
        {
            if ( x > b.global_hi[d] )
                *xp = x - b.extent[d];
            else if ( x < b.global_lo[d] )
                *xp = x + b.extent[d];
        }
        else
            drop = drop || outside;
    }
    keep[p] = drop ? 0 : 1;
    count_flagged( drop, dropped );
}

// Ascending list of the kept particles: idx[rank[p]] = p for every kept p,
// rank being the exclusive scan of the keep flags (the Serial order of the
// particles that stay, Cabana_Distributor.hpp:80-101).
__global__ __launch_bounds__( BLOCK ) void compact_index_kernel( const int32_t* keep,
                                                                 const int64_t* rank, int64_t n,
                                                                 int64_t* idx )
{
    const int64_t p = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if ( p < n && keep[p] )
        idx[rank[p]] = p;
}

// Cell of a particle for the library's own binning of a call that passes
// none: the floor of its logical cell coordinate in the owned cells,
// clamped into them (a particle outside them, or with a NaN coordinate,
// still gets a cell; the tiled kernels interpolate a particle whose stencil
// leaves its brick through global memory).  Key = cardinal cell index.
__global__ __launch_bounds__( BLOCK ) void auto_bin_key_kernel( PFieldDev pos, int64_t n,
                                                                MeshDev m, int32_t* keys )
{
#pragma clang fp contract( off )
    const int64_t p = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if ( p >= n )
        return;
    const uint32_t b = pos.base( p );
    int c[3];
#pragma unroll
    for ( int d = 0; d < 3; ++d )
    {
        const double x = *pos.ptr<const double>( b, d );
        const double f = floor( ( x - m.low[d] ) * m.rdx[d] ) - (double)m.halo;
        c[d] = f >= 0.0 ? ( f < (double)m.num_cell[d] ? (int)f : m.num_cell[d] - 1 ) : 0;
    }
    keys[p] = ( c[0] * m.num_cell[1] + c[1] ) * m.num_cell[2] + c[2];
}

// ---------------------------------------------------------------------------
// BOV output
// ---------------------------------------------------------------------------

// The written block of a grid array in file order, x fastest after the
// dofs: out[((k * n1 + j) * n0 + i) * dofs + l] = grid(lo + i, lo + j,
// lo + k, l) (reorderView and the C-order subarray of
// Cabana_Grid_BovWriter.hpp:82-152).
template <class T>
__global__ __launch_bounds__( BLOCK ) void bov_pack_kernel( GFieldDev g, int lo, int n0, int n1,
                                                            int n2, T* out )
{
    const int dofs = g.ext[3];
    const int64_t total = (int64_t)n0 * n1 * n2 * dofs;
    const int64_t t = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if ( t >= total )
        return;
    int64_t r = t;
    const int l = (int)( r % dofs );
    r /= dofs;
    const int i = (int)( r % n0 );
    r /= n0;
    const int j = (int)( r % n1 );
    const int k = (int)( r / n1 );
    out[t] = *gptr<const T>( g, ( lo + i ) * g.s[0] + ( lo + j ) * g.s[1] + ( lo + k ) * g.s[2] +
                                    l * g.s[3] );
}

} // namespace

hipError_t launch_halo_gather( Launcher& L, bool fp64, const MeshDev& m,
                               const GFieldDev& g, int width, hipStream_t s )
{
    const int64_t n = (int64_t)g.ext[0] * g.ext[1] * g.ext[2];
    L.before( "halo_gather", s );
    if ( fp64 )
        hipLaunchKernelGGL( halo_fill_kernel<double>, blocks_for( n ),
                            dim3( BLOCK ), 0, s, m, g, width, 0 );
    else
        hipLaunchKernelGGL( halo_fill_kernel<float>, blocks_for( n ),
                            dim3( BLOCK ), 0, s, m, g, width, 0 );
    L.after( "halo_gather", s );
    return hipGetLastError();
}

hipError_t launch_halo_adjoint_fill( Launcher& L, bool fp64, const MeshDev& m,
                                     const GFieldDev& g, int width,
                                     hipStream_t s )
{
    const int64_t n = (int64_t)g.ext[0] * g.ext[1] * g.ext[2];
    L.before( "halo_adjoint_fill", s );
    if ( fp64 )
        hipLaunchKernelGGL( halo_fill_kernel<double>, blocks_for( n ),
                            dim3( BLOCK ), 0, s, m, g, width, 1 );
    else
        hipLaunchKernelGGL( halo_fill_kernel<float>, blocks_for( n ),
                            dim3( BLOCK ), 0, s, m, g, width, 1 );
    L.after( "halo_adjoint_fill", s );
    return hipGetLastError();
}

hipError_t launch_halo_scatter_sum( Launcher& L, bool fp64, const MeshDev& m,
                                    const GFieldDev& g, int width,
                                    hipStream_t s )
{
    int64_t n = 1;
    for ( int d = 0; d < 3; ++d )
        n *= m.num_cell[d] + ( m.periodic[d] ? 0 : 1 );
    L.before( "halo_scatter_sum", s );
    if ( fp64 )
        hipLaunchKernelGGL( halo_fold_kernel<double>, blocks_for( n ),
                            dim3( BLOCK ), 0, s, m, g, width );
    else
        hipLaunchKernelGGL( halo_fold_kernel<float>, blocks_for( n ),
                            dim3( BLOCK ), 0, s, m, g, width );
    L.after( "halo_scatter_sum", s );
    return hipGetLastError();
}

hipError_t launch_zero_ghosts( Launcher& L, bool fp64, const MeshDev& m,
                               const GFieldDev& g, hipStream_t s )
{
    const int64_t n = (int64_t)g.ext[0] * g.ext[1] * g.ext[2];
    L.before( "zero_ghosts", s );
    if ( fp64 )
        hipLaunchKernelGGL( zero_ghosts_kernel<double>, blocks_for( n ),
                            dim3( BLOCK ), 0, s, m, g );
    else
        hipLaunchKernelGGL( zero_ghosts_kernel<float>, blocks_for( n ),
                            dim3( BLOCK ), 0, s, m, g );
    L.after( "zero_ghosts", s );
    return hipGetLastError();
}

hipError_t launch_lcl_keys( Launcher& L, const PFieldDev& pos, int64_t begin,
                            int64_t end, const CartGridDev& g, int32_t bad_key,
                            int32_t* keys, unsigned long long* bad,
                            hipStream_t s )
{
    if ( end <= begin )
        return hipSuccess;
    L.before( "lcl_keys", s );
    hipLaunchKernelGGL( lcl_key_kernel, blocks_for( end - begin ),
                        dim3( BLOCK ), 0, s, pos, begin, end, g, bad_key, keys,
                        bad );
    L.after( "lcl_keys", s );
    return hipGetLastError();
}

int64_t lcl_sort_tiles( int64_t n )
{
    return ( n + RS_TILE - 1 ) / RS_TILE;
}

int lcl_sort_passes( int key_bits, int* digit_bits )
{
    const int passes = key_bits <= 0 ? 1 : ( key_bits + RS_MAX_BITS - 1 ) / RS_MAX_BITS;
    *digit_bits = key_bits <= 0 ? 1 : ( key_bits + passes - 1 ) / passes;
    return passes;
}

hipError_t launch_lcl_sort( Launcher& L, int64_t n, int key_bits,
                            int32_t* keys_a, int32_t* keys_b, int32_t* vals_a,
                            int32_t* vals_b, int32_t* hist, int64_t* digit_base,
                            int64_t* block_sums, int64_t nblock_cap,
                            int64_t* permutes, int64_t begin, int32_t** sorted,
                            hipStream_t s )
{
    *sorted = keys_a;
    if ( n <= 0 )
        return hipSuccess;
    int dbits;
    const int passes = lcl_sort_passes( key_bits, &dbits );
    const int nbins = 1 << dbits;
    const int64_t ntiles = lcl_sort_tiles( n );
    int32_t* kin = keys_a;
    int32_t* kout = keys_b;
    int32_t* vin = vals_a;
    int32_t* vout = vals_b;
    for ( int p = 0; p < passes; ++p )
    {
        const int shift = p * dbits;
        L.before( "lcl_sort_hist", s );
        hipLaunchKernelGGL( radix_hist_kernel, dim3( (unsigned)ntiles ),
                            dim3( RS_BLOCK ), 0, s, kin, n, shift, nbins,
                            ntiles, hist );
        L.after( "lcl_sort_hist", s );
        hipError_t e = launch_exclusive_scan( L, hist, digit_base,
                                              (int64_t)nbins * ntiles,
                                              block_sums, nblock_cap, s );
        if ( e != hipSuccess )
            return e;
        const bool first = p == 0, last = p == passes - 1;
        L.before( "lcl_sort_scatter", s );
        if ( first && last )
            hipLaunchKernelGGL( ( radix_scatter_kernel<true, true> ),
                                dim3( (unsigned)ntiles ), dim3( RS_BLOCK ), 0,
                                s, kin, vin, n, shift, nbins, ntiles,
                                digit_base, kout, vout, permutes, begin );
        else if ( first )
            hipLaunchKernelGGL( ( radix_scatter_kernel<true, false> ),
                                dim3( (unsigned)ntiles ), dim3( RS_BLOCK ), 0,
                                s, kin, vin, n, shift, nbins, ntiles,
                                digit_base, kout, vout, permutes, begin );
        else if ( last )
            hipLaunchKernelGGL( ( radix_scatter_kernel<false, true> ),
                                dim3( (unsigned)ntiles ), dim3( RS_BLOCK ), 0,
                                s, kin, vin, n, shift, nbins, ntiles,
                                digit_base, kout, vout, permutes, begin );
        else
            hipLaunchKernelGGL( ( radix_scatter_kernel<false, false> ),
                                dim3( (unsigned)ntiles ), dim3( RS_BLOCK ), 0,
                                s, kin, vin, n, shift, nbins, ntiles,
                                digit_base, kout, vout, permutes, begin );
        L.after( "lcl_sort_scatter", s );
        e = hipGetLastError();
        if ( e != hipSuccess )
            return e;
        int32_t* t = kin;
        kin = kout;
        kout = t;
        t = vin;
        vin = vout;
        vout = t;
    }
    *sorted = kin;
    return hipSuccess;
}

hipError_t launch_lcl_bounds( Launcher& L, const int32_t* sorted_keys,
                              int64_t n, int64_t ncell, int32_t* counts,
                              int64_t* offsets, hipStream_t s )
{
    if ( ncell <= 0 )
        return hipSuccess;
    L.before( "lcl_bounds", s );
    hipLaunchKernelGGL( lcl_bounds_kernel, blocks_for( ncell ), dim3( BLOCK ),
                        0, s, sorted_keys, n, ncell, counts, offsets );
    L.after( "lcl_bounds", s );
    return hipGetLastError();
}

hipError_t launch_exclusive_scan( Launcher& L, const int32_t* in, int64_t* out,
                                  int64_t n, int64_t* block_sums,
                                  int64_t nblock_cap, hipStream_t s )
{
    if ( n <= 0 )
        return hipSuccess;
    const int64_t nb = ( n + SCAN_BLOCK - 1 ) / SCAN_BLOCK;
    if ( nb > nblock_cap )
        return hipErrorInvalidValue;
    L.before( "exclusive_scan", s );
    hipLaunchKernelGGL( scan_reduce_kernel, dim3( (unsigned)nb ),
                        dim3( SCAN_BLOCK ), 0, s, in, n, block_sums );
    hipLaunchKernelGGL( scan_block_sums_kernel, dim3( 1 ), dim3( SCAN_BLOCK ),
                        0, s, block_sums, nb );
    hipLaunchKernelGGL( scan_apply_kernel, dim3( (unsigned)nb ),
                        dim3( SCAN_BLOCK ), 0, s, in, out, n, block_sums );
    L.after( "exclusive_scan", s );
    return hipGetLastError();
}
namespace
{
// Words per particle of a row-contiguous field, or 0 when the field's
// components are interleaved with other particles' (vector length > 1).
int row_field_words( const PFieldDev& f, int elem_size, int64_t* row_words )
{
    if ( f.vl != 1 || ( f.ncomp * elem_size ) % 4 != 0 || ( f.stride * elem_size ) % 4 != 0 ||
         reinterpret_cast<uintptr_t>( f.data ) % 4 != 0 )
        return 0;
    *row_words = (int64_t)f.stride * elem_size / 4;
    return f.ncomp * elem_size / 4;
}

#define CA_NW_CASES( KER, ... )                                                       \
    switch ( nw )                                                                     \
    {                                                                                 \
    case 1: hipLaunchKernelGGL( KER<1>, grid, dim3( BLOCK ), 0, s, __VA_ARGS__ ); break;   \
    case 2: hipLaunchKernelGGL( KER<2>, grid, dim3( BLOCK ), 0, s, __VA_ARGS__ ); break;   \
    case 3: hipLaunchKernelGGL( KER<3>, grid, dim3( BLOCK ), 0, s, __VA_ARGS__ ); break;   \
    case 4: hipLaunchKernelGGL( KER<4>, grid, dim3( BLOCK ), 0, s, __VA_ARGS__ ); break;   \
    case 6: hipLaunchKernelGGL( KER<6>, grid, dim3( BLOCK ), 0, s, __VA_ARGS__ ); break;   \
    case 8: hipLaunchKernelGGL( KER<8>, grid, dim3( BLOCK ), 0, s, __VA_ARGS__ ); break;   \
    case 9: hipLaunchKernelGGL( KER<9>, grid, dim3( BLOCK ), 0, s, __VA_ARGS__ ); break;   \
    case 12: hipLaunchKernelGGL( KER<12>, grid, dim3( BLOCK ), 0, s, __VA_ARGS__ ); break; \
    case 18: hipLaunchKernelGGL( KER<18>, grid, dim3( BLOCK ), 0, s, __VA_ARGS__ ); break; \
    default: hipLaunchKernelGGL( KER<0>, grid, dim3( BLOCK ), 0, s, __VA_ARGS__ ); break;  \
    }

void launch_gather_words( int nw, int64_t nwords, hipStream_t s, const uint32_t* src,
                          int64_t row_words, const int64_t* permutes, uint32_t* scratch )
{
    const dim3 grid( (unsigned)( ( nwords + PERMUTE_WPT * BLOCK - 1 ) / ( PERMUTE_WPT * BLOCK ) ) );
    CA_NW_CASES( permute_gather_words_kernel, src, row_words, nw, permutes, nwords, scratch )
}

void launch_put_words( int nw, int64_t nwords, hipStream_t s, uint32_t* dst, int64_t row_words,
                       int64_t begin, const uint32_t* scratch )
{
    const dim3 grid = blocks_for( nwords );
    CA_NW_CASES( permute_put_words_kernel, dst, row_words, nw, begin, nwords, scratch )
}
#undef CA_NW_CASES
} // namespace

hipError_t launch_permute_gather( Launcher& L, const PFieldDev& f,
                                  int elem_size, const int64_t* permutes,
                                  int64_t begin, int64_t end, void* scratch,
                                  hipStream_t s )
{
    if ( end <= begin )
        return hipSuccess;
    L.before( "permute_gather", s );
    int64_t row_words = 0;
    const int nwg = row_field_words( f, elem_size, &row_words );
    if ( nwg && ( end - begin ) * nwg < ( int64_t( 1 ) << 31 ) )
    {
        const int nw = nwg;
        const int64_t nwords = ( end - begin ) * nw;
        launch_gather_words( nw, nwords, s, reinterpret_cast<const uint32_t*>( f.data ),
                             row_words, permutes, reinterpret_cast<uint32_t*>( scratch ) );
        L.after( "permute_gather", s );
        return hipGetLastError();
    }
    if ( elem_size == 8 )
        hipLaunchKernelGGL( permute_gather_kernel<uint64_t>,
                            blocks_for( end - begin ), dim3( BLOCK ), 0, s, f,
                            permutes, begin, end,
                            reinterpret_cast<uint64_t*>( scratch ) );
    else
        hipLaunchKernelGGL( permute_gather_kernel<uint32_t>,
                            blocks_for( end - begin ), dim3( BLOCK ), 0, s, f,
                            permutes, begin, end,
                            reinterpret_cast<uint32_t*>( scratch ) );
    L.after( "permute_gather", s );
    return hipGetLastError();
}

hipError_t launch_permute_copy_back( Launcher& L, const PFieldDev& f,
                                     int elem_size, int64_t begin,
                                     int64_t end, const void* scratch,
                                     hipStream_t s )
{
    if ( end <= begin )
        return hipSuccess;
    L.before( "permute_copy_back", s );
    int64_t row_words = 0;
    const int nwp = row_field_words( f, elem_size, &row_words );
    if ( nwp && ( end - begin ) * nwp < ( int64_t( 1 ) << 31 ) )
    {
        const int nw = nwp;
        const int64_t nwords = ( end - begin ) * nw;
        uint32_t* dst = reinterpret_cast<uint32_t*>( f.data );
        hipError_t e = hipSuccess;
        if ( row_words == nw )
            e = hipMemcpyAsync( dst + begin * row_words, scratch, (size_t)nwords * 4,
                                hipMemcpyDeviceToDevice, s );
        else
            launch_put_words( nw, nwords, s, dst, row_words, begin,
                              reinterpret_cast<const uint32_t*>( scratch ) );
        L.after( "permute_copy_back", s );
        return e != hipSuccess ? e : hipGetLastError();
    }
    if ( elem_size == 8 )
        hipLaunchKernelGGL( permute_copy_back_kernel<uint64_t>,
                            blocks_for( end - begin ), dim3( BLOCK ), 0, s, f,
                            begin, end,
                            reinterpret_cast<const uint64_t*>( scratch ) );
    else
        hipLaunchKernelGGL( permute_copy_back_kernel<uint32_t>,
                            blocks_for( end - begin ), dim3( BLOCK ), 0, s, f,
                            begin, end,
                            reinterpret_cast<const uint32_t*>( scratch ) );
    L.after( "permute_copy_back", s );
    return hipGetLastError();
}

hipError_t launch_permute_members( Launcher& L, const PFieldDev* f,
                                   const int* elem_size, int count,
                                   const int64_t* permutes, int64_t begin,
                                   int64_t end, void* scratch, hipStream_t s )
{
    const int64_t n = end - begin;
    if ( n <= 0 || count <= 0 )
        return hipSuccess;
    // Members that fit the fused pass: row-contiguous, at most
    // PERMUTE_MAX_MEMBERS of them, each with n * words below 2^31.
    PermuteMembers pm{};
    int index[PERMUTE_MAX_MEMBERS];
    int64_t words = 0;
    for ( int k = 0; k < count && pm.count < PERMUTE_MAX_MEMBERS; ++k )
    {
        int64_t rw = 0;
        const int nw = row_field_words( f[k], elem_size[k], &rw );
        if ( !nw || n * nw >= ( int64_t( 1 ) << 31 ) )
            continue;
        const int m = pm.count++;
        index[m] = k;
        pm.src[m] = reinterpret_cast<const uint32_t*>( f[k].data );
        pm.row_words[m] = rw;
        pm.words[m] = nw;
        pm.word_offset[m] = words;
        words += nw;
    }
    uint32_t* scr = reinterpret_cast<uint32_t*>( scratch );
    if ( pm.count >= 2 )
    {
        L.before( "permute_gather", s );
        const dim3 grid( (unsigned)( ( n + PERMUTE_CHUNK - 1 ) / PERMUTE_CHUNK ) );
        hipLaunchKernelGGL( permute_gather_members_kernel, grid, dim3( BLOCK ), 0, s, pm,
                            permutes, n, scr );
        L.after( "permute_gather", s );
        hipError_t e = hipGetLastError();
        if ( e != hipSuccess )
            return e;
        for ( int m = 0; m < pm.count; ++m )
        {
            L.before( "permute_copy_back", s );
            const int nw = pm.words[m];
            const int64_t rw = pm.row_words[m];
            uint32_t* dst = reinterpret_cast<uint32_t*>( f[index[m]].data );
            const uint32_t* blk = scr + n * pm.word_offset[m];
            if ( rw == nw )
                e = hipMemcpyAsync( dst + begin * rw, blk, (size_t)( n * nw ) * 4,
                                    hipMemcpyDeviceToDevice, s );
            else
                launch_put_words( nw, n * nw, s, dst, rw, begin, blk );
            L.after( "permute_copy_back", s );
            if ( e == hipSuccess )
                e = hipGetLastError();
            if ( e != hipSuccess )
                return e;
        }
    }
    // The other members one at a time (the scratch is reused in stream order).
    for ( int k = 0; k < count; ++k )
    {
        bool done = false;
        for ( int m = 0; m < pm.count && pm.count >= 2; ++m )
            done = done || index[m] == k;
        if ( done )
            continue;
        hipError_t e = launch_permute_gather( L, f[k], elem_size[k], permutes, begin, end,
                                              scratch, s );
        if ( e == hipSuccess )
            e = launch_permute_copy_back( L, f[k], elem_size[k], begin, end, scratch, s );
        if ( e != hipSuccess )
            return e;
    }
    return hipSuccess;
}

hipError_t launch_fill_i32( Launcher& L, int32_t* dst, int64_t n,
                            int32_t value, hipStream_t s )
{
    if ( n <= 0 )
        return hipSuccess;
    L.before( "fill_i32", s );
    hipLaunchKernelGGL( fill_i32_kernel, blocks_for( n ), dim3( BLOCK ), 0, s,
                        dst, n, value );
    L.after( "fill_i32", s );
    return hipGetLastError();
}
hipError_t launch_brick_parts( Launcher& L, const TileArgs& t,
                               const int32_t num_cell[3], int32_t* brick_nparts,
                               int64_t* brick_start, int64_t* block_sums,
                               int64_t nblock_cap, int2* parts,
                               int32_t* num_parts, hipStream_t s )
{
    const int64_t nbricks = (int64_t)t.ntile[0] * t.ntile[1] * t.ntile[2];
    if ( nbricks <= 0 )
        return hipErrorInvalidValue;
    L.before( "brick_parts", s );
    hipLaunchKernelGGL( brick_count_kernel, blocks_for( nbricks ), dim3( BLOCK ), 0, s,
                        t, make_int3( num_cell[0], num_cell[1], num_cell[2] ),
                        brick_nparts );
    L.after( "brick_parts", s );
    hipError_t e = launch_exclusive_scan( L, brick_nparts, brick_start, nbricks,
                                          block_sums, nblock_cap, s );
    if ( e != hipSuccess )
        return e;
    L.before( "brick_parts", s );
    hipLaunchKernelGGL( part_fill_kernel, blocks_for( nbricks ), dim3( BLOCK ), 0, s,
                        brick_nparts, brick_start, (int)nbricks, parts, num_parts );
    L.after( "brick_parts", s );
    return hipGetLastError();
}

hipError_t launch_migrate_count( Launcher& L, const PFieldDev& pos, int64_t n,
                                 const MigrateBoundsDev& b, unsigned long long* count,
                                 hipStream_t s )
{
    if ( n <= 0 )
        return hipSuccess;
    L.before( "migrate_count", s );
    hipLaunchKernelGGL( migrate_count_kernel, blocks_for( n ), dim3( BLOCK ), 0, s, pos, n, b,
                        count );
    L.after( "migrate_count", s );
    return hipGetLastError();
}

hipError_t launch_migrate_shift( Launcher& L, const PFieldDev& pos, int64_t n,
                                 const MigrateBoundsDev& b, int32_t* keep,
                                 unsigned long long* dropped, hipStream_t s )
{
    if ( n <= 0 )
        return hipSuccess;
    L.before( "migrate_shift", s );
    hipLaunchKernelGGL( migrate_shift_kernel, blocks_for( n ), dim3( BLOCK ), 0, s, pos, n, b,
                        keep, dropped );
    L.after( "migrate_shift", s );
    return hipGetLastError();
}

hipError_t launch_compact_index( Launcher& L, const int32_t* keep, const int64_t* rank,
                                 int64_t n, int64_t* idx, hipStream_t s )
{
    if ( n <= 0 )
        return hipSuccess;
    L.before( "compact_index", s );
    hipLaunchKernelGGL( compact_index_kernel, blocks_for( n ), dim3( BLOCK ), 0, s, keep, rank, n,
                        idx );
    L.after( "compact_index", s );
    return hipGetLastError();
}

hipError_t launch_bov_pack( Launcher& L, bool fp64, const GFieldDev& g, int lo, const int n[3],
                            void* out, hipStream_t s )
{
    const int64_t total = (int64_t)n[0] * n[1] * n[2] * g.ext[3];
    if ( total <= 0 )
        return hipSuccess;
    L.before( "bov_pack", s );
    if ( fp64 )
        hipLaunchKernelGGL( bov_pack_kernel<double>, blocks_for( total ), dim3( BLOCK ), 0, s, g,
                            lo, n[0], n[1], n[2], static_cast<double*>( out ) );
    else
        hipLaunchKernelGGL( bov_pack_kernel<float>, blocks_for( total ), dim3( BLOCK ), 0, s, g,
                            lo, n[0], n[1], n[2], static_cast<float*>( out ) );
    L.after( "bov_pack", s );
    return hipGetLastError();
}

hipError_t launch_auto_bin_keys( Launcher& L, const PFieldDev& pos, int64_t n, const MeshDev& m,
                                 int32_t* keys, hipStream_t s )
{
    if ( n <= 0 )
        return hipSuccess;
    L.before( "auto_bin_keys", s );
    hipLaunchKernelGGL( auto_bin_key_kernel, blocks_for( n ), dim3( BLOCK ), 0, s, pos, n, m,
                        keys );
    L.after( "auto_bin_keys", s );
    return hipGetLastError();
}

} // namespace cabana_amd
