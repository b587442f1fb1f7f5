// Internal types shared by the C-ABI layer and the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/cabana_amd.h"

namespace cabana_amd
{

// Mesh constants as the kernels see them.  low/rdx follow evaluateSpline
// (Cabana_Grid_Splines.hpp:1283-1342): low = coordinate of ghosted node 0,
// dx = coordinate(node 1) - coordinate(node 0), rdx = 1 / dx.
struct MeshDev
{
    double low[3];
    double rdx[3];
    int32_t num_node[3]; // ghosted node extent
    int32_t num_cell[3];
    int32_t halo;        // halo cell width of the local grid
    int32_t periodic[3];
    int32_t stag[3];     // 1 where the arrays' entity sits at cell centres
};

// Particle field addressing (Cabana_Slice.hpp:119-190).  For a power-of-two
// vector length the block and lane come from shift/mask; when one block holds
// every particle, shift = 62 so the block index is always 0.
// Particle field (Cabana slice addressing, Cabana_Slice.hpp:119-190):
// element (p, c) at data[(p / vl) * stride + p % vl + vl * c].  The vector
// length is a power of two (as every Cabana AoSoA's) or covers all particles,
// so the index is a shift and a mask; every element index of the field fits
// in 32 bits.  Both are checked on the host; device addressing is 32-bit
// integer math plus one 64-bit pointer add, with no branches.
struct PFieldDev
{
    char* data;
    uint32_t vl;
    uint32_t stride;
    int32_t shift;
    int32_t ncomp;
    uint32_t mask;

    __device__ __forceinline__ uint32_t base( int64_t p64 ) const
    {
        const uint32_t p = (uint32_t)p64;
        return ( p >> shift ) * stride + ( p & mask );
    }
    template <class T>
    __device__ __forceinline__ T* ptr( uint32_t b, int c ) const
    {
        return reinterpret_cast<T*>( data + (uint64_t)( b + vl * (uint32_t)c ) * sizeof( T ) );
    }
    // Components [0, N) of particle p.  CONTIG (vector_length 1, checked
    // by the caller's host code) moves them as 16/8-byte vectors (4-byte
    // aligned access), one instruction per 16 bytes instead of one per
    // component; otherwise (AoSoA) each component is its own access,
    // coalesced across lanes.
    template <class T, int N, bool CONTIG = false>
    __device__ __forceinline__ void load( int64_t p, T* v ) const;
    template <class T, int N, bool CONTIG = false>
    __device__ __forceinline__ void store( int64_t p, const T* v ) const;
};

namespace detail
{
template <class T, int W>
struct VecOf;
template <>
struct VecOf<float, 4>
{
    typedef float type __attribute__( ( ext_vector_type( 4 ), aligned( 4 ) ) );
};
template <>
struct VecOf<float, 2>
{
    typedef float type __attribute__( ( ext_vector_type( 2 ), aligned( 4 ) ) );
};
template <>
struct VecOf<double, 2>
{
    typedef double type __attribute__( ( ext_vector_type( 2 ), aligned( 8 ) ) );
};

template <class T, int N>
__device__ __forceinline__ void vec_copy_in( const T* src, T* dst )
{
    constexpr int W = 16 / (int)sizeof( T );
    int i = 0;
#pragma unroll
    for ( ; i + W <= N; i += W )
    {
        const typename VecOf<T, W>::type x = *reinterpret_cast<const typename VecOf<T, W>::type*>( src + i );
#pragma unroll
        for ( int e = 0; e < W; ++e )
            dst[i + e] = x[e];
    }
    if constexpr ( W == 4 && N % 4 >= 2 )
    {
        const typename VecOf<T, 2>::type x = *reinterpret_cast<const typename VecOf<T, 2>::type*>( src + i );
        dst[i] = x[0];
        dst[i + 1] = x[1];
        i += 2;
    }
#pragma unroll
    for ( ; i < N; ++i )
        dst[i] = src[i];
}

template <class T, int N>
__device__ __forceinline__ void vec_copy_out( const T* src, T* dst )
{
    constexpr int W = 16 / (int)sizeof( T );
    int i = 0;
#pragma unroll
    for ( ; i + W <= N; i += W )
    {
        typename VecOf<T, W>::type x;
#pragma unroll
        for ( int e = 0; e < W; ++e )
            x[e] = src[i + e];
        *reinterpret_cast<typename VecOf<T, W>::type*>( dst + i ) = x;
    }
    if constexpr ( W == 4 && N % 4 >= 2 )
    {
        typename VecOf<T, 2>::type x;
        x[0] = src[i];
        x[1] = src[i + 1];
        *reinterpret_cast<typename VecOf<T, 2>::type*>( dst + i ) = x;
        i += 2;
    }
#pragma unroll
    for ( ; i < N; ++i )
        dst[i] = src[i];
}
} // namespace detail

template <class T, int N, bool CONTIG>
__device__ __forceinline__ void PFieldDev::load( int64_t p, T* v ) const
{
    const uint32_t b = base( p );
    if constexpr ( CONTIG )
    {
        detail::vec_copy_in<T, N>( ptr<const T>( b, 0 ), v );
    }
    else
    {
#pragma unroll
        for ( int c = 0; c < N; ++c )
            v[c] = *ptr<const T>( b, c );
    }
}

template <class T, int N, bool CONTIG>
__device__ __forceinline__ void PFieldDev::store( int64_t p, const T* v ) const
{
    const uint32_t b = base( p );
    if constexpr ( CONTIG )
    {
        detail::vec_copy_out<T, N>( v, ptr<T>( b, 0 ) );
    }
    else
    {
#pragma unroll
        for ( int c = 0; c < N; ++c )
            *ptr<T>( b, c ) = v[c];
    }
}

// Grid array: element strides are 32-bit and every element's byte offset fits
// in 32 bits (checked on the host), so device addressing is a uniform 64-bit
// base plus a 32-bit offset.
struct GFieldDev
{
    char* data;
    int32_t s[4];
    int32_t ext[4];
};

template <class T>
__device__ __forceinline__ T* gptr( const GFieldDev& g, int elem )
{
    return reinterpret_cast<T*>(
        g.data + (uint32_t)( (uint32_t)elem * (uint32_t)sizeof( T ) ) );
}

// Bilinear forms B(P, G; W, dW) between a particle vector P and a node vector
// G.  P2G scatters dB/dG, G2P gathers dB/dP and the position adjoint
// differentiates B through the weights.
enum Form : int32_t
{
    FORM_SV = 0,  // P[1]  G[1]: P0 G0 W
    FORM_VV = 1,  // P[3]  G[3]: sum_d Pd Gd W
    FORM_SG = 2,  // P[1]  G[3]: sum_d P0 Gd dW_d
    FORM_VD = 3,  // P[3]  G[1]: sum_d Pd G0 dW_d
    FORM_TD = 4,  // P[9]  G[3]: sum P[3a+b] G[a] dW_b
    FORM_TDT = 5, // P[9]  G[3]: sum P[3a+b] G[b] dW_a
    FORM_COUNT = 6
};

// Forward operator -> form.  P2G: ScalarValue->SV, VectorValue->VV,
// ScalarGradient->SG, VectorDivergence->VD, TensorDivergence->TD.
// G2P: ScalarValue->SV, VectorValue->VV, ScalarGradient->VD,
// VectorGradient->TDT, VectorDivergence->SG.
inline int p2g_form( int op )
{
    static const int f[5] = { FORM_SV, FORM_VV, FORM_SG, FORM_VD, FORM_TD };
    return f[op];
}
inline int g2p_form( int op )
{
    static const int f[5] = { FORM_SV, FORM_VV, FORM_VD, FORM_TDT, FORM_SG };
    return f[op];
}
inline int form_pcomp( int form )
{
    static const int n[6] = { 1, 3, 1, 3, 9, 9 };
    return n[form];
}
inline int form_gdof( int form )
{
    static const int n[6] = { 1, 3, 3, 1, 3, 3 };
    return n[form];
}

constexpr int MAX_TERMS = 3;       // terms fused into one scatter pass
constexpr int MAX_GATHER_TERMS = 5; // terms fused into one gather pass
constexpr int MAX_CALL_TERMS = 16; // terms accepted by one API call

// One term of a scatter pass (P2G forward, or G2P adjoint):
//   grid += mult * dB/dG (P = pin)
//   xbar += mult * dB/dx (P = pin, G = gx)     [only when the pass has xbar]
struct ScatterTermDev
{
    PFieldDev pin;
    GFieldDev grid;
    GFieldDev gx;
    double mult;
    int32_t form;
};

// One term of a gather pass (G2P forward, or P2G adjoint):
//   pout += mult * dB/dP (G = grid)
//   xbar += mult * dB/dx (P = px, G = grid)     [only when the pass has xbar]
struct GatherTermDev
{
    PFieldDev pout;
    PFieldDev px;
    GFieldDev grid;
    double mult;
    int32_t form;
};

struct ScatterArgs
{
    MeshDev mesh;
    PFieldDev points;
    PFieldDev xbar;
    int64_t begin;
    int64_t end;
    int32_t nterms;
    int32_t has_xbar;
    int32_t contig; // every particle field of the pass has vector_length 1
    ScatterTermDev t[MAX_TERMS];
    unsigned long long* skipped;
};

struct GatherArgs
{
    MeshDev mesh;
    PFieldDev points;
    PFieldDev xbar;
    int64_t begin;
    int64_t end;
    int32_t nterms;
    int32_t has_xbar;
    int32_t contig; // every particle field of the pass has vector_length 1
    GatherTermDev t[MAX_GATHER_TERMS];
    unsigned long long* skipped;
};

// Every term of a call in one pass over the particles, for any set of
// operators: the spline is evaluated once per particle, then the G2P terms
// run in call order (each adds into its particle field) and after them the
// P2G terms (each reads its particle field, which a G2P term of the same
// particle may just have updated, and adds into its grid with atomics).
// One term of such a pass: its particle field (output of a G2P term, input
// of a P2G term), grid, multiplier and form.
struct MultiTermDev
{
    PFieldDev f;
    GFieldDev grid;
    double mult;
    int32_t form;
};
struct MultiTermArgs
{
    MeshDev mesh;
    PFieldDev points;
    int64_t begin;
    int64_t end;
    int32_t n_g2p;
    int32_t n_p2g;
    MultiTermDev g[MAX_CALL_TERMS];
    MultiTermDev s[MAX_CALL_TERMS];
    unsigned long long* skipped;
};
// Kernel arguments are limited to 4 KiB.
static_assert( sizeof( MultiTermArgs ) <= 4000, "MultiTermArgs exceeds the kernel argument space" );
inline MultiTermDev multi_term( const GatherTermDev& t ) { return { t.pout, t.grid, t.mult, t.form }; }
inline MultiTermDev multi_term( const ScatterTermDev& t ) { return { t.pin, t.grid, t.mult, t.form }; }

// Deterministic P2G of one term: every node of the term's grid sums the
// contributions of the particles of the cells around it, cell by cell in
// cardinal order and in binned order inside a cell, and adds the sum to
// itself (one writer per node, no atomics).  The binning covers every
// particle (cell keys clamped into the owned cells) and lists them through
// perm.
struct PullArgs
{
    MeshDev mesh;
    PFieldDev points;
    // One term, or the MPM set (ScalarValue, VectorValue, TensorDivergence)
    // in one pass; every term's grid has the same extents.
    MultiTermDev term[3];
    int32_t nterms;
    const int32_t* counts;
    const int64_t* offsets;
    const int64_t* perm;
};

// Brick tiling of cell-sorted particles.  A brick is tile[0] x tile[1] x
// tile[2] mesh cells; its particles are the tile[0] * tile[1] cell runs along
// the last dimension (contiguous in the cell-sorted order).  Bricks are cut
// into parts of at most `cap` particles; a workgroup processes one part at a
// time from the part table (filled on the device before the launch).
struct TileArgs
{
    const int64_t* cell_offsets; // exclusive scan, per mesh cell
    const int32_t* cell_counts;
    const int2* parts;           // (brick, part index within the brick)
    const int32_t* num_parts;    // number of entries of `parts`
    int32_t tile[3];             // cells per brick and dimension
    int32_t ntile[3];            // bricks per dimension
    int32_t win[3];              // window nodes per dimension
    int32_t flush_perm[3];       // grid dims from largest to smallest stride
    int32_t max_runs;            // tile[0] * tile[1]
    int32_t max_win_nodes;       // win[0] * win[1] * win[2]
    int32_t max_base_cells;      // stencil-base cells of a brick
    int32_t cap;                 // particles per part
    int32_t max_tasks;           // scatter: task-table capacity
    int32_t grid;                // workgroups launched (a multiple of 8)
    int64_t bin_begin;           // particle index of cell offset 0
    // Particle at each binned position q (perm[q - bin_begin]) when the
    // binning does not reorder the particles themselves; nullptr: the
    // particle at q is q.
    const int64_t* perm;
    __device__ __forceinline__ int64_t particle( int64_t q ) const
    {
        return perm ? perm[q - bin_begin] : q;
    }
};

// LDS record of one particle in the cell-sorted scatter: 3 fractional
// offsets then the premultiplied attributes, padded to 16 bytes so a record
// is read with 128-bit LDS loads.  Records are read through a sorted index,
// so no stride is free of bank conflicts and none is padded for it.
__host__ __device__ constexpr int record_stride( int len, int elem_bytes )
{
    return ( len * elem_bytes + 15 ) / 16 * 16 / elem_bytes;
}
// Threads of a scatter workgroup (FP32 / FP64 records) and the particles
// each thread holds per part; their product is the largest part.
constexpr int scatter_block( bool fp64 ) { return 256; }
constexpr int SCATTER_SLOTS = 2; // tested value; 3 failed the unbinned tests
// Longest run of one base cell's records that one scatter task takes
// (8 measured faster than 16 or 32 at 16 particles per cell: tasks of more
// even length leave fewer lanes idle).
constexpr int SCATTER_TASK_MAX = 8; // tested value; 12 failed the dense-cell tests
constexpr int scatter_cap_max( bool fp64 ) { return scatter_block( fp64 ) * SCATTER_SLOTS; }

// Launch-time error/timing context owned by the C-ABI layer.
class Launcher
{
  public:
    virtual ~Launcher() = default;
    // Called around every kernel launch; name is the kernel family.
    virtual void before( const char* name, hipStream_t s ) = 0;
    virtual void after( const char* name, hipStream_t s ) = 0;
};

// Fused term sets with their own kernels; any other multi-term call runs one
// single-form pass per term.
enum Plan : int32_t
{
    PLAN_SINGLE = 0,  // one term, any form
    PLAN_MPM_P2G = 1, // (SV, VV, TD): mass, momentum, stress divergence
    PLAN_MPM_G2P = 2, // (VV, TDT) on one shared grid: velocity + gradient
    PLAN_MASS_MOMENTUM = 3, // (SV, VV): mass and momentum (P2G only)
    // Gather only: the P2G adjoint of (SV, VV, TD) with the position part
    // of the G2P adjoint of (VV, TDT) on one shared grid, in one pass.
    PLAN_MPM_ADJOINT = 4
};

// Kernel launchers (return hipSuccess or the launch error), one set per
// spline order (compiled in separate translation units).
#define CABANA_AMD_DECLARE_LAUNCHERS( SUFFIX )                                 \
    hipError_t launch_gather_##SUFFIX( Launcher& L, bool fp64, int plan,       \
                                       const GatherArgs& a, hipStream_t s );   \
    hipError_t launch_scatter_atomic_##SUFFIX( Launcher& L, bool fp64,         \
                                               int plan, const ScatterArgs& a, \
                                               hipStream_t s );                \
    hipError_t launch_multi_term_##SUFFIX( Launcher& L, bool fp64,             \
                                           const MultiTermArgs& a, hipStream_t s ); \
    hipError_t launch_p2g_pull_##SUFFIX( Launcher& L, bool fp64, const PullArgs& a, \
                                         hipStream_t s );                      \
    hipError_t launch_count_outside_##SUFFIX( Launcher& L, const MeshDev& m,   \
                                              const PFieldDev& pts, int64_t n, \
                                              unsigned long long* skipped,     \
                                              hipStream_t s );
#define CABANA_AMD_DECLARE_TILED( SUFFIX )                                     \
    hipError_t launch_gather_tiled_##SUFFIX( Launcher& L, bool fp64, int plan, \
                                             const GatherArgs& a,              \
                                             const TileArgs& t, size_t lds,    \
                                             hipStream_t s );                  \
    hipError_t launch_scatter_cells_##SUFFIX( Launcher& L, bool fp64,          \
                                              int plan, const ScatterArgs& a,  \
                                              const TileArgs& t, size_t lds,   \
                                              hipStream_t s );
CABANA_AMD_DECLARE_TILED( o1 )
CABANA_AMD_DECLARE_TILED( o2 )
CABANA_AMD_DECLARE_TILED( o3 )
CABANA_AMD_DECLARE_LAUNCHERS( o1 )
CABANA_AMD_DECLARE_LAUNCHERS( o2 )
CABANA_AMD_DECLARE_LAUNCHERS( o3 )
// Orders 0 and 4 to 6 run the per-particle (untiled) kernels only.
CABANA_AMD_DECLARE_LAUNCHERS( o0 )
CABANA_AMD_DECLARE_LAUNCHERS( o4 )
CABANA_AMD_DECLARE_LAUNCHERS( o5 )
CABANA_AMD_DECLARE_LAUNCHERS( o6 )

// Halo kernels on a ghosted node grid.
hipError_t launch_halo_gather( Launcher& L, bool fp64, const MeshDev& m,
                               const GFieldDev& g, int width, hipStream_t s );
hipError_t launch_halo_scatter_sum( Launcher& L, bool fp64, const MeshDev& m,
                                    const GFieldDev& g, int width,
                                    hipStream_t s );
// Transpose of the fold for a cotangent grid: ghost := owner within the band,
// 0 for every other non-owned node.
hipError_t launch_halo_adjoint_fill( Launcher& L, bool fp64, const MeshDev& m,
                                     const GFieldDev& g, int width,
                                     hipStream_t s );
// Zero every non-owned node.
hipError_t launch_zero_ghosts( Launcher& L, bool fp64, const MeshDev& m,
                               const GFieldDev& g, hipStream_t s );

// Binning.
struct CartGridDev
{
    double min[3];
    double rdx[3];
    int32_t nx[3];
};
// Bin keys (cell index, or bad_key outside the grid) of [begin, end).
hipError_t launch_lcl_keys( Launcher& L, const PFieldDev& pos, int64_t begin,
                            int64_t end, const CartGridDev& g, int32_t bad_key,
                            int32_t* keys, unsigned long long* bad,
                            hipStream_t s );
hipError_t launch_exclusive_scan( Launcher& L, const int32_t* in, int64_t* out,
                                  int64_t n, int64_t* block_sums,
                                  int64_t nblock_cap, hipStream_t s );
// Stable radix sort of n keys in keys_a (values: the indices 0..n-1), digit
// passes over key_bits bits.  Every pass needs hist (int32) and digit_base
// (int64) of lcl_sort_bins * lcl_sort_tiles entries.  The sorted values
// leave as begin + index in permutes; *sorted names the buffer holding the
// sorted keys (keys_a or keys_b).
int64_t lcl_sort_tiles( int64_t n );
int lcl_sort_passes( int key_bits, int* digit_bits );
hipError_t launch_lcl_sort( Launcher& L, int64_t n, int key_bits,
                            int32_t* keys_a, int32_t* keys_b, int32_t* vals_a,
                            int32_t* vals_b, int32_t* hist, int64_t* digit_base,
                            int64_t* block_sums, int64_t nblock_cap,
                            int64_t* permutes, int64_t begin, int32_t** sorted,
                            hipStream_t s );
// counts[c] and offsets[c] of bins 0..ncell-1 from n sorted keys.
hipError_t launch_lcl_bounds( Launcher& L, const int32_t* sorted_keys,
                              int64_t n, int64_t ncell, int32_t* counts,
                              int64_t* offsets, hipStream_t s );
// Row-contiguous members of one particle set, permuted in one fused pass:
// member f's rows are row_words[f] 4-byte words apart, words[f] of them per
// particle, and its block of the scratch starts at n * word_offset[f] words.
constexpr int PERMUTE_MAX_MEMBERS = 16;
struct PermuteMembers
{
    int32_t count;
    int32_t words[PERMUTE_MAX_MEMBERS];
    const uint32_t* src[PERMUTE_MAX_MEMBERS];
    int64_t row_words[PERMUTE_MAX_MEMBERS];
    int64_t word_offset[PERMUTE_MAX_MEMBERS];
};
// Permutes count members in place (field[i] = old field[permutes[i - begin]]
// for i in [begin, end)); scratch holds at least
// sum over members of (elem_size * ncomp) * (end - begin + 1) bytes.
hipError_t launch_permute_members( Launcher& L, const PFieldDev* f,
                                   const int* elem_size, int count,
                                   const int64_t* permutes, int64_t begin,
                                   int64_t end, void* scratch, hipStream_t s );
hipError_t launch_permute_gather( Launcher& L, const PFieldDev& f,
                                  int elem_size, const int64_t* permutes,
                                  int64_t begin, int64_t end, void* scratch,
                                  hipStream_t s );
hipError_t launch_permute_copy_back( Launcher& L, const PFieldDev& f,
                                     int elem_size, int64_t begin,
                                     int64_t end, const void* scratch,
                                     hipStream_t s );
hipError_t launch_fill_i32( Launcher& L, int32_t* dst, int64_t n,
                            int32_t value, hipStream_t s );

// Bounds of particleMigrate on one rank (Cabana_Grid_ParticleDistributor.hpp
// :87-148, :170-221), formed on the host as LocalMesh and GlobalMesh form
// them: the band of migrateCount (ghosted corners moved in by
// min_halo_width cells), the owned corners, the global corners and extent.
struct MigrateBoundsDev
{
    double band_lo[3];
    double band_hi[3];
    double own_lo[3];
    double own_hi[3];
    double global_lo[3];
    double global_hi[3];
    double extent[3];
    int32_t periodic[3];
};
// *count += particles outside the band.
hipError_t launch_migrate_count( Launcher& L, const PFieldDev& pos, int64_t n,
                                 const MigrateBoundsDev& b, unsigned long long* count,
                                 hipStream_t s );
// Periodic shift of the positions in place, keep[p] = 0 for particles
// leaving through a non-periodic boundary, *dropped += their number.
hipError_t launch_migrate_shift( Launcher& L, const PFieldDev& pos, int64_t n,
                                 const MigrateBoundsDev& b, int32_t* keep,
                                 unsigned long long* dropped, hipStream_t s );
// Owned-cell keys of particles [0, n) for the library's own binning
// (clamped into the owned cells).
hipError_t launch_auto_bin_keys( Launcher& L, const PFieldDev& pos, int64_t n, const MeshDev& m,
                                 int32_t* keys, hipStream_t s );
// Block [lo, lo + n) of every dimension of a grid array, packed x fastest
// after the dofs (the BOV data order) into out.
hipError_t launch_bov_pack( Launcher& L, bool fp64, const GFieldDev& g, int lo, const int n[3],
                            void* out, hipStream_t s );
// idx[rank[p]] = p for the kept particles (rank: exclusive scan of keep).
hipError_t launch_compact_index( Launcher& L, const int32_t* keep, const int64_t* rank,
                                 int64_t n, int64_t* idx, hipStream_t s );
// Part table of a brick tiling: brick_nparts (nbricks ints), brick_start
// (nbricks int64), block_sums (scan scratch, nblock_cap int64), parts
// (nbricks + n / cap entries at most) and num_parts (1 int).
hipError_t launch_brick_parts( Launcher& L, const TileArgs& t,
                               const int32_t num_cell[3], int32_t* brick_nparts,
                               int64_t* brick_start, int64_t* block_sums,
                               int64_t nblock_cap, int2* parts,
                               int32_t* num_parts, hipStream_t s );

// p2g with the brick part table of a tiled scatter built in `scratch`
// (scratch_bytes long) instead of the context's allocations; the context's
// device must be current.  Records an error on the context and returns its
// status.
cabana_amd_status p2g_fixed_scratch( cabana_amd_context* ctx, const cabana_amd_mesh* mesh,
                                     int32_t order, const cabana_amd_particle_field* points,
                                     int32_t num_terms, const cabana_amd_p2g_term* terms,
                                     int32_t halo_width, const cabana_amd_binning* bins,
                                     cabana_amd_stream stream, void* scratch,
                                     size_t scratch_bytes );
// Records msg as the context's (and the thread's) last error; returns st.
cabana_amd_status fail_with( cabana_amd_context* ctx, cabana_amd_status st, const char* msg );

// LDS atomic-add throughput probe (ca_probe.hip): iters adds per lane over
// num_groups workgroups of 256 threads; ms = kernel time.
hipError_t probe_lds_add( bool fp64, int iters, int num_groups, void* out, hipStream_t s,
                          float& ms );

} // namespace cabana_amd
