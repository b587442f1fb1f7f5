// C-ABI layer: descriptor validation, plan selection and kernel dispatch.
#include "ca_internal.hpp"

#include <rocprofiler-sdk-roctx/roctx.h>

#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

using namespace cabana_amd;

namespace
{

thread_local std::string g_thread_error = "";

// Binning serials issued so far (cabana_amd_new_binning_serial).
std::atomic<uint64_t> g_serials{ 0 };

// Build generation of every counts/offsets array that
// cabana_amd_linked_cell_build wrote, across all contexts: a part table
// records the generations of the arrays it was derived from and is reused
// only while both are unchanged.
std::mutex g_gen_mutex;
std::map<const void*, uint64_t> g_generation;
uint64_t g_generation_clock = 0;

uint64_t generation_of( const void* p )
{
    std::lock_guard<std::mutex> lock( g_gen_mutex );
    const auto it = g_generation.find( p );
    return it == g_generation.end() ? 0 : it->second;
}

void bump_generation( const void* a, const void* b )
{
    std::lock_guard<std::mutex> lock( g_gen_mutex );
    const uint64_t g = ++g_generation_clock;
    g_generation[a] = g;
    g_generation[b] = g;
}

struct Pending
{
    std::string name;
    hipEvent_t start;
    hipEvent_t stop;
};

struct Stat
{
    int64_t launches = 0;
    double ms = 0.0;
};

} // namespace

struct cabana_amd_context : public Launcher
{
    int device = 0;
    std::string error;
    unsigned long long* skipped = nullptr; // device counter
    unsigned long long* bad = nullptr;     // device counter (binning)

    // Reusable scratch.
    static constexpr int NUM_SLOTS = 7;
    void* scratch[NUM_SLOTS] = {};
    size_t scratch_bytes[NUM_SLOTS] = {};

    // Brick part tables of binnings that carry a build serial, each in its
    // own allocation (the general scratch slots move when they grow); the
    // least recently used one is rebuilt when a new key arrives.
    struct PartTable
    {
        const void* counts = nullptr;
        const void* offsets = nullptr;
        int64_t begin = 0;
        int64_t end = 0;
        uint64_t serial = 0; // 0: empty
        uint64_t gen_counts = 0;  // build generations of counts/offsets
        uint64_t gen_offsets = 0; // when the table was made
        int32_t tile[3] = { 0, 0, 0 };
        int32_t cap = 0;
        void* mem = nullptr;
        size_t bytes = 0;
        int2* parts = nullptr;
        int32_t* num_parts = nullptr;
        uint64_t used = 0;
    };
    static constexpr int PART_TABLES = 8;
    PartTable tables[PART_TABLES];
    uint64_t table_clock = 0;

    // The cell binning the library builds itself for calls that pass none
    // (see auto_bins): counts, offsets and the particle of each binned
    // position, in SLOT_BINS.
    bool auto_binning = true;
    // P2G through the deterministic pull kernel (cabana_amd_set_deterministic).
    bool deterministic = false;
    cabana_amd_binning auto_bins = {};
    const int64_t* auto_perm = nullptr;

    // Kernel timing.
    bool profiling = false;
    std::vector<Pending> pending;
    std::vector<hipEvent_t> pool;
    std::map<std::string, Stat> stats;
    std::vector<std::string> order;

    hipEvent_t get_event()
    {
        if ( !pool.empty() )
        {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        hipEventCreate( &e );
        return e;
    }

    void before( const char* name, hipStream_t s ) override
    {
        if ( !profiling )
            return;
        Pending p;
        p.name = name;
        p.start = get_event();
        p.stop = nullptr;
        hipEventRecord( p.start, s );
        pending.push_back( p );
    }

    void after( const char* name, hipStream_t s ) override
    {
        if ( !profiling || pending.empty() )
            return;
        Pending& p = pending.back();
        if ( p.name != name || p.stop )
            return;
        p.stop = get_event();
        hipEventRecord( p.stop, s );
    }

    hipError_t drain()
    {
        hipError_t first = hipSuccess;
        for ( auto& p : pending )
        {
            if ( p.stop )
            {
                hipError_t e = hipEventSynchronize( p.stop );
                float ms = 0.0f;
                if ( e == hipSuccess )
                    e = hipEventElapsedTime( &ms, p.start, p.stop );
                if ( e != hipSuccess && first == hipSuccess )
                    first = e;
                if ( !stats.count( p.name ) )
                    order.push_back( p.name );
                Stat& st = stats[p.name];
                st.launches += 1;
                st.ms += ms;
                pool.push_back( p.stop );
            }
            pool.push_back( p.start );
        }
        pending.clear();
        return first;
    }

    hipError_t ensure( int slot, size_t bytes )
    {
        if ( scratch_bytes[slot] >= bytes )
            return hipSuccess;
        if ( scratch[slot] )
            hipFree( scratch[slot] );
        scratch[slot] = nullptr;
        scratch_bytes[slot] = 0;
        size_t want = bytes + bytes / 4 + 256;
        hipError_t e = hipMalloc( &scratch[slot], want );
        if ( e == hipSuccess )
            scratch_bytes[slot] = want;
        return e;
    }
};

namespace
{

// roctx range over one entry point, named as the reference's Kokkos
// profiling regions (Cabana_Grid_Halo_Mpi.hpp:80,170,
// Cabana_LinkedCellList.hpp:655, Cabana_Sort.hpp:557,608) where it has them;
// a no-op unless a tracing tool is attached.
struct TraceRange
{
    explicit TraceRange( const char* name ) { roctxRangePushA( name ); }
    ~TraceRange() { roctxRangePop(); }
    TraceRange( const TraceRange& ) = delete;
    TraceRange& operator=( const TraceRange& ) = delete;
};

enum Slot
{
    SLOT_CELL_OF = 0,
    SLOT_CURSOR = 1,
    SLOT_SCAN = 2,
    SLOT_PERMUTE = 3,
    SLOT_TILES = 4,
    SLOT_SNAPSHOT = 5, // copy of a grid that one pass both reads and adds into
    SLOT_BINS = 6      // the library's own cell binning of an unbinned call
};

cabana_amd_status fail( cabana_amd_context* ctx, cabana_amd_status st,
                        const std::string& msg )
{
    if ( ctx )
        ctx->error = msg;
    g_thread_error = msg;
    return st;
}

cabana_amd_status hip_fail( cabana_amd_context* ctx, hipError_t e,
                            const char* where )
{
    std::string msg = std::string( where ) + ": " + hipGetErrorString( e );
    return fail( ctx, e == hipErrorOutOfMemory ? CABANA_AMD_ERROR_OUT_OF_MEMORY
                                               : CABANA_AMD_ERROR_DEVICE,
                 msg );
}

#define CA_HIP( ctx, expr, where )                                            \
    do                                                                         \
    {                                                                          \
        hipError_t e__ = ( expr );                                             \
        if ( e__ != hipSuccess )                                               \
            return hip_fail( ctx, e__, where );                                \
    } while ( 0 )

#define CA_CHECK( st )                                                         \
    do                                                                         \
    {                                                                          \
        cabana_amd_status s__ = ( st );                                        \
        if ( s__ != CABANA_AMD_OK )                                            \
            return s__;                                                        \
    } while ( 0 )

int dtype_size( int dtype )
{
    switch ( dtype )
    {
    case CABANA_AMD_FLOAT64:
    case CABANA_AMD_INT64:
        return 8;
    case CABANA_AMD_FLOAT32:
    case CABANA_AMD_INT32:
        return 4;
    default:
        return 0;
    }
}

// GlobalMesh + LocalMesh + evaluateSpline constants
// (Cabana_Grid_GlobalMesh.hpp:59-80,160-165; Cabana_Grid_LocalMesh.hpp:89-104;
// Cabana_Grid_Splines.hpp:1296-1318).
cabana_amd_status make_mesh( cabana_amd_context* ctx,
                             const cabana_amd_mesh* mesh, MeshDev& m )
{
    if ( !mesh )
        return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT, "mesh is NULL" );
    if ( mesh->halo_cell_width < 0 )
        return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT,
                     "negative halo cell width" );
    if ( mesh->entity < CABANA_AMD_NODE || mesh->entity > CABANA_AMD_EDGE_K )
        return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT, "unknown mesh entity" );
    for ( int d = 0; d < 3; ++d )
    {
        const double cell = mesh->cell_size[d];
        const double extent = mesh->high_corner[d] - mesh->low_corner[d];
        if ( !( cell > 0.0 ) || !( extent > 0.0 ) )
            return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT,
                         "mesh extent and cell size must be positive" );
        const double nc = std::rint( extent / cell );
        if ( nc < 1.0 || nc > 1.0e8 )
            return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT,
                         "mesh cell count out of range" );
        const int n = (int)nc;
        if ( std::fabs( n * cell - extent ) > 100.0 * 2.220446049250313e-16 )
            return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT,
                         "Extent not evenly divisible by uniform cell size" );
        // Entity per dimension (Cabana_Grid_LocalMesh.hpp:176-237): Cell at
        // cell centres, Face<D> at nodes only in D, Edge<D> at centres only
        // in D, Node at nodes.
        const int e = mesh->entity;
        const int stag = e == CABANA_AMD_NODE   ? 0
                         : e == CABANA_AMD_CELL ? 1
                         : e <= CABANA_AMD_FACE_K ? ( d == e - CABANA_AMD_FACE_I ? 0 : 1 )
                                                  : ( d == e - CABANA_AMD_EDGE_I ? 1 : 0 );
        m.stag[d] = stag;
        m.num_cell[d] = n;
        m.num_node[d] = n + 2 * mesh->halo_cell_width + 1 - stag;
        m.periodic[d] = mesh->periodic[d] ? 1 : 0;
        // evaluateSpline (Cabana_Grid_Splines.hpp:1283-1318): low corner =
        // coordinates of entity 0, dx = coordinates of entity 1 minus it.
        const double own_low = mesh->low_corner[d] + cell * 0;
        const double ghost_low = own_low - mesh->halo_cell_width * cell;
        const double x0 = stag ? ghost_low + ( 0.0 + 0.5 ) * cell : ghost_low + (double)0 * cell;
        const double x1 = stag ? ghost_low + ( 1.0 + 0.5 ) * cell : ghost_low + (double)1 * cell;
        m.low[d] = x0;
        m.rdx[d] = 1.0 / ( x1 - x0 );
    }
    m.halo = mesh->halo_cell_width;
    return CABANA_AMD_OK;
}

cabana_amd_status make_pfield( cabana_amd_context* ctx,
                               const cabana_amd_particle_field& f,
                               int ncomp, int dtype, int64_t n_needed,
                               bool allow_null, const char* what,
                               PFieldDev& out )
{
    std::memset( &out, 0, sizeof( out ) );
    if ( !f.data )
    {
        if ( allow_null )
            return CABANA_AMD_OK;
        return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT,
                     std::string( what ) + ": data is NULL" );
    }
    if ( ncomp > 0 && f.num_comp != ncomp )
        return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT,
                     std::string( what ) + ": expected " +
                         std::to_string( ncomp ) + " components, got " +
                         std::to_string( f.num_comp ) );
    if ( f.num_comp < 1 )
        return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT,
                     std::string( what ) + ": no components" );
    if ( dtype >= 0 && f.dtype != dtype )
        return fail( ctx, CABANA_AMD_ERROR_UNSUPPORTED,
                     std::string( what ) + ": unsupported dtype" );
    if ( f.num_particle < n_needed )
        return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT,
                     std::string( what ) + ": fewer particles than points" );
    if ( f.vector_length < 1 )
        return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT,
                     std::string( what ) + ": vector_length < 1" );
    if ( f.num_particle > f.vector_length &&
         f.soa_stride < f.vector_length * f.num_comp )
        return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT,
                     std::string( what ) +
                         ": soa_stride smaller than vector_length * num_comp" );
    // Device addressing is 32-bit: particle indices below 2^31 and every
    // element index of the field below 2^32.
    {
        const int64_t np = f.num_particle;
        const int64_t vl = f.vector_length;
        const int64_t last = np > 0 ? ( ( np - 1 ) / vl ) * f.soa_stride +
                                          ( np - 1 ) % vl + vl * ( f.num_comp - 1 )
                                    : vl * ( f.num_comp - 1 );
        if ( np >= ( int64_t( 1 ) << 31 ) || f.soa_stride < 0 || last >= ( int64_t( 1 ) << 32 ) ||
             vl >= ( int64_t( 1 ) << 32 ) )
            return fail( ctx, CABANA_AMD_ERROR_UNSUPPORTED,
                         std::string( what ) +
                             ": more than 2^31 particles or 2^32 elements in one field" );
    }
    out.data = static_cast<char*>( f.data );
    out.vl = (uint32_t)f.vector_length;
    out.stride = (uint32_t)f.soa_stride;
    out.ncomp = f.num_comp;
    if ( f.num_comp == 1 && f.vector_length >= f.num_particle )
    {
        // A single-component field inside one vector: element p sits at p,
        // the same address as vector_length 1 with unit stride.
        out.vl = 1;
        out.stride = 1;
        out.shift = 0;
        out.mask = 0;
    }
    else if ( f.vector_length >= f.num_particle )
    {
        out.shift = 31;
        out.mask = 0x7fffffffu;
    }
    else if ( ( f.vector_length & ( f.vector_length - 1 ) ) == 0 )
    {
        int sh = 0;
        while ( ( int64_t( 1 ) << sh ) < f.vector_length )
            ++sh;
        out.shift = sh;
        out.mask = (uint32_t)( f.vector_length - 1 );
    }
    else
    {
        return fail( ctx, CABANA_AMD_ERROR_UNSUPPORTED,
                     std::string( what ) +
                         ": vector_length must be a power of two (or cover every particle)" );
    }
    return CABANA_AMD_OK;
}

cabana_amd_status make_gfield( cabana_amd_context* ctx, const MeshDev& m,
                               const cabana_amd_grid_field& g, int dofs,
                               int dtype, bool allow_null, const char* what,
                               GFieldDev& out )
{
    std::memset( &out, 0, sizeof( out ) );
    if ( !g.data )
    {
        if ( allow_null )
            return CABANA_AMD_OK;
        return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT,
                     std::string( what ) + ": data is NULL" );
    }
    for ( int d = 0; d < 3; ++d )
        if ( g.extent[d] != m.num_node[d] )
            return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT,
                         std::string( what ) +
                             ": extent does not match the ghosted node space "
                             "(num_cell + 2 * halo + 1)" );
    if ( dofs > 0 && g.extent[3] != dofs )
        return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT,
                     std::string( what ) + ": expected " +
                         std::to_string( dofs ) + " dofs, got " +
                         std::to_string( g.extent[3] ) );
    if ( g.extent[3] < 1 )
        return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT,
                     std::string( what ) + ": no dofs" );
    if ( dtype >= 0 && g.dtype != dtype )
        return fail( ctx, CABANA_AMD_ERROR_UNSUPPORTED,
                     std::string( what ) + ": unsupported dtype" );
    for ( int d = 0; d < 4; ++d )
        if ( g.stride[d] < 1 && g.extent[d] > 1 )
            return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT,
                         std::string( what ) + ": non-positive stride" );
    // Device addressing uses 32-bit element strides and byte offsets.
    int64_t max_elem = 0;
    for ( int d = 0; d < 4; ++d )
        max_elem += ( g.extent[d] - 1 ) * g.stride[d];
    const int64_t es = ( g.dtype == CABANA_AMD_FLOAT64 || g.dtype == CABANA_AMD_INT64 ) ? 8 : 4;
    if ( max_elem >= ( int64_t( 1 ) << 31 ) ||
         ( max_elem + 1 ) * es > ( int64_t( 1 ) << 32 ) )
        return fail( ctx, CABANA_AMD_ERROR_UNSUPPORTED,
                     std::string( what ) + ": grid arrays above 4 GiB are not supported" );
    out.data = static_cast<char*>( g.data );
    for ( int d = 0; d < 4; ++d )
    {
        out.s[d] = (int32_t)g.stride[d];
        out.ext[d] = g.extent[d];
    }
    return CABANA_AMD_OK;
}

bool same_grid( const GFieldDev& a, const GFieldDev& b )
{
    return a.data == b.data && a.s[0] == b.s[0] && a.s[1] == b.s[1] &&
           a.s[2] == b.s[2] && a.s[3] == b.s[3];
}

cabana_amd_status check_order( cabana_amd_context* ctx, int order )
{
    if ( order < 0 || order > 6 )
        return fail( ctx, CABANA_AMD_ERROR_UNSUPPORTED,
                     "spline order must be 0 to 6" );
    return CABANA_AMD_OK;
}

int resolve_width( const MeshDev& m, int w ) { return w < 0 ? m.halo : w; }

cabana_amd_status check_width( cabana_amd_context* ctx, const MeshDev& m,
                               int w )
{
    if ( w > m.halo )
        return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT,
                     "Requested halo width larger than local grid halo" );
    return CABANA_AMD_OK;
}

// Particle precision from the first term: FLOAT64 plain or FLOAT32 mixed.
cabana_amd_status precision_of( cabana_amd_context* ctx, int dtype,
                                bool& fp64 )
{
    if ( dtype == CABANA_AMD_FLOAT64 )
        fp64 = true;
    else if ( dtype == CABANA_AMD_FLOAT32 )
        fp64 = false;
    else
        return fail( ctx, CABANA_AMD_ERROR_UNSUPPORTED,
                     "values and grids must be FLOAT64 or FLOAT32" );
    return CABANA_AMD_OK;
}

hipError_t run_gather( cabana_amd_context* ctx, int order, bool fp64,
                       int plan, const GatherArgs& a, hipStream_t s )
{
    switch ( order )
    {
    case 0:
        return launch_gather_o0( *ctx, fp64, plan, a, s );
    case 1:
        return launch_gather_o1( *ctx, fp64, plan, a, s );
    case 2:
        return launch_gather_o2( *ctx, fp64, plan, a, s );
    case 3:
        return launch_gather_o3( *ctx, fp64, plan, a, s );
    case 4:
        return launch_gather_o4( *ctx, fp64, plan, a, s );
    case 5:
        return launch_gather_o5( *ctx, fp64, plan, a, s );
    default:
        return launch_gather_o6( *ctx, fp64, plan, a, s );
    }
}

hipError_t run_scatter_atomic( cabana_amd_context* ctx, int order, bool fp64,
                               int plan, const ScatterArgs& a, hipStream_t s )
{
    switch ( order )
    {
    case 0:
        return launch_scatter_atomic_o0( *ctx, fp64, plan, a, s );
    case 1:
        return launch_scatter_atomic_o1( *ctx, fp64, plan, a, s );
    case 2:
        return launch_scatter_atomic_o2( *ctx, fp64, plan, a, s );
    case 3:
        return launch_scatter_atomic_o3( *ctx, fp64, plan, a, s );
    case 4:
        return launch_scatter_atomic_o4( *ctx, fp64, plan, a, s );
    case 5:
        return launch_scatter_atomic_o5( *ctx, fp64, plan, a, s );
    default:
        return launch_scatter_atomic_o6( *ctx, fp64, plan, a, s );
    }
}

hipError_t run_gather_tiled( cabana_amd_context* ctx, int order, bool fp64,
                             int plan, const GatherArgs& a, const TileArgs& t,
                             size_t lds, hipStream_t s )
{
    switch ( order )
    {
    case 1:
        return launch_gather_tiled_o1( *ctx, fp64, plan, a, t, lds, s );
    case 2:
        return launch_gather_tiled_o2( *ctx, fp64, plan, a, t, lds, s );
    default:
        return launch_gather_tiled_o3( *ctx, fp64, plan, a, t, lds, s );
    }
}

hipError_t run_scatter_cells( cabana_amd_context* ctx, int order, bool fp64,
                              int plan, const ScatterArgs& a,
                              const TileArgs& t, size_t lds, hipStream_t s )
{
    switch ( order )
    {
    case 1:
        return launch_scatter_cells_o1( *ctx, fp64, plan, a, t, lds, s );
    case 2:
        return launch_scatter_cells_o2( *ctx, fp64, plan, a, t, lds, s );
    default:
        return launch_scatter_cells_o3( *ctx, fp64, plan, a, t, lds, s );
    }
}

hipError_t run_multi_term( cabana_amd_context* ctx, int order, bool fp64,
                          const MultiTermArgs& a, hipStream_t s )
{
    switch ( order )
    {
    case 0:
        return launch_multi_term_o0( *ctx, fp64, a, s );
    case 1:
        return launch_multi_term_o1( *ctx, fp64, a, s );
    case 2:
        return launch_multi_term_o2( *ctx, fp64, a, s );
    case 3:
        return launch_multi_term_o3( *ctx, fp64, a, s );
    case 4:
        return launch_multi_term_o4( *ctx, fp64, a, s );
    case 5:
        return launch_multi_term_o5( *ctx, fp64, a, s );
    default:
        return launch_multi_term_o6( *ctx, fp64, a, s );
    }
}

hipError_t run_pull( cabana_amd_context* ctx, int order, bool fp64, const PullArgs& a,
                     hipStream_t s )
{
    switch ( order )
    {
    case 0:
        return launch_p2g_pull_o0( *ctx, fp64, a, s );
    case 1:
        return launch_p2g_pull_o1( *ctx, fp64, a, s );
    case 2:
        return launch_p2g_pull_o2( *ctx, fp64, a, s );
    case 3:
        return launch_p2g_pull_o3( *ctx, fp64, a, s );
    case 4:
        return launch_p2g_pull_o4( *ctx, fp64, a, s );
    case 5:
        return launch_p2g_pull_o5( *ctx, fp64, a, s );
    default:
        return launch_p2g_pull_o6( *ctx, fp64, a, s );
    }
}

hipError_t run_count_outside( cabana_amd_context* ctx, int order, const MeshDev& m,
                              const PFieldDev& pts, int64_t n, hipStream_t s )
{
    switch ( order )
    {
    case 0:
        return launch_count_outside_o0( *ctx, m, pts, n, ctx->skipped, s );
    case 1:
        return launch_count_outside_o1( *ctx, m, pts, n, ctx->skipped, s );
    case 2:
        return launch_count_outside_o2( *ctx, m, pts, n, ctx->skipped, s );
    case 3:
        return launch_count_outside_o3( *ctx, m, pts, n, ctx->skipped, s );
    case 4:
        return launch_count_outside_o4( *ctx, m, pts, n, ctx->skipped, s );
    case 5:
        return launch_count_outside_o5( *ctx, m, pts, n, ctx->skipped, s );
    default:
        return launch_count_outside_o6( *ctx, m, pts, n, ctx->skipped, s );
    }
}

// Whether a pass over these particles can run the cell-tiled kernels (Node
// arrays only).
bool tiled_call( const cabana_amd_binning* bins, int order, bool node_arrays )
{
    return bins && bins->counts && bins->offsets && order >= 1 && order <= 3 && node_arrays;
}

bool node_arrays( const MeshDev& m ) { return !m.stag[0] && !m.stag[1] && !m.stag[2]; }

// Index of the first term of a form, or -1.
template <class Term>
int find_form( const Term* T, int n, int form, const std::vector<bool>& taken )
{
    for ( int i = 0; i < n; ++i )
        if ( !taken[i] && T[i].form == form )
            return i;
    return -1;
}

size_t a16( size_t b ) { return ( b + 15 ) & ~(size_t)15; }

enum TileKind
{
    TILE_GATHER = 0,
    TILE_SCATTER = 1
};

constexpr int TILE_BLOCK = 256;

// Workgroups of a tiled launch (a multiple of 8, one range per XCD): 4096 by
// default; CABANA_AMD_MAX_TILE_GROUPS lowers it (tests use it to give every
// workgroup many parts).
int64_t max_tile_groups()
{
    const char* e = std::getenv( "CABANA_AMD_MAX_TILE_GROUPS" );
    long x = e ? std::strtol( e, nullptr, 10 ) : 0;
    if ( x < 8 || x > 4096 )
        x = 4096;
    return (int64_t)( x / 8 * 8 );
}

struct TileShape
{
    int T[3];
    int cap;
    int max_tasks;
};

int base_cells( const int T[3], int order )
{
    const int spread = ( order == 2 ) ? 1 : 0;
    return ( T[0] + spread ) * ( T[1] + spread ) * ( T[2] + spread );
}

// LDS bytes of the tiled kernels; mirrors their in-kernel layout.
size_t tile_lds_bytes( TileKind kind, bool fp64, const int* slot_ng, int nslot,
                       int rec_len, int order, const TileShape& sh )
{
    const int extra = ( order == 1 ) ? 1 : 3;
    const size_t wn = (size_t)( sh.T[0] + extra ) * ( sh.T[1] + extra ) * ( sh.T[2] + extra );
    const size_t runs = (size_t)sh.T[0] * sh.T[1];
    const size_t real = fp64 ? 8 : 4;
    size_t b = 0;
    if ( kind == TILE_GATHER )
    {
        for ( int i = 0; i < nslot; ++i )
            b = a16( b + real * wn * ( slot_ng[i] == 1 ? 1 : 4 ) );
        b = a16( b + 8 * runs );
        b = a16( b + 4 * ( runs + 1 ) );
        return b;
    }
    int dt = 0;
    for ( int i = 0; i < nslot; ++i )
        dt += slot_ng[i];
    const size_t nbase = (size_t)base_cells( sh.T, order );
    const size_t rs = (size_t)record_stride( rec_len, (int)real );
    b = a16( 8 * (size_t)dt * wn );
    b = a16( b + 4 * ( nbase + 1 ) );
    b = a16( b + 4 * (size_t)sh.max_tasks ); // sorted task table
    b = a16( b + 4 * (size_t)sh.max_tasks ); // unsorted tasks
    b = a16( b + 4 * (size_t)sh.max_tasks ); // their length bucket and rank
    b = a16( b + 4 * 2 * SCATTER_TASK_MAX ); // length histogram and offsets
    b = a16( b + 2 * (size_t)sh.cap );
    b = a16( b + real * rs * (size_t)sh.cap );
    for ( int buf = 0; buf < 2; ++buf )
    {
        b = a16( b + 8 * runs );
        b = a16( b + 4 * ( runs + 1 ) );
    }
    b = a16( b + 8 * 128 ); // part-table entries
    b += 4 * 16 + 16;       // per-wave scan sums of a workgroup of up to 1024 threads
    return b;
}

// Brick shape and part capacity.  Gathers use the largest cubic brick whose
// windows fit 52 KB of LDS for FP32 (three workgroups per CU) and 78 KB for
// FP64 (whose registers allow two), with parts of ~1.25x the mean brick
// population.  Scatters sort a whole part in LDS: the brick is the largest
// near-cubic one whose expected population (x1.2) fits a part (at most two
// particles per thread) that, with the FP64 window, stays within 78 KB of
// LDS (two workgroups per CU run at once, the kernel's register budget);
// denser bricks are cut into several parts.
bool choose_shape( TileKind kind, bool fp64, const int* slot_ng, int nslot, int rec_len,
                   int order, double mean_ppc, TileShape& out )
{
    if ( kind == TILE_GATHER )
    {
        static const int cand[] = { 8, 7, 6, 5, 4, 3, 2 };
        for ( int c : cand )
        {
            TileShape sh = { { c, c, c }, 0, 0 };
            if ( tile_lds_bytes( kind, fp64, slot_ng, nslot, rec_len, order, sh ) <=
                 ( fp64 ? 78 : 52 ) * 1024 )
            {
                const double mean = mean_ppc * c * c * c;
                double cap = 1.25 * mean;
                cap = cap < 4096 ? 4096 : ( cap > ( 1 << 20 ) ? ( 1 << 20 ) : cap );
                sh.cap = ( (int)cap + 255 ) / 256 * 256;
                out = sh;
                return true;
            }
        }
        return false;
    }
    const size_t budget = 78 * 1024;
    auto with_cap = [&]( TileShape sh, int cap ) {
        sh.cap = cap;
        sh.max_tasks = base_cells( sh.T, order ) + cap / SCATTER_TASK_MAX + 1;
        return sh;
    };
    bool found = false;
    TileShape best = {};
    int best_vol = 0;
    for ( int a = 2; a <= 8; ++a )
        for ( int b = a; b <= a + 1 && b <= 8; ++b )
            for ( int c = b; c <= b + 1 && c <= 8; ++c )
            {
                const int vol = a * b * c;
                if ( vol <= best_vol )
                    continue;
                int cap = (int)( 1.2 * mean_ppc * vol ) + 32;
                cap = ( cap + 63 ) / 64 * 64;
                cap = cap < 256 ? 256 : cap;
                if ( cap > scatter_cap_max( fp64 ) )
                    continue;
                TileShape sh = with_cap( { { a, b, c }, 0, 0 }, cap );
                if ( tile_lds_bytes( kind, fp64, slot_ng, nslot, rec_len, order, sh ) <=
                     budget )
                {
                    best = sh;
                    best_vol = vol;
                    found = true;
                }
            }
    if ( !found )
    {
        // Dense: the smallest brick with the largest part that fits.
        for ( int cap = scatter_cap_max( fp64 ); cap >= 64; cap -= 64 )
        {
            TileShape sh = with_cap( { { 2, 2, 2 }, 0, 0 }, cap );
            if ( tile_lds_bytes( kind, fp64, slot_ng, nslot, rec_len, order, sh ) <= budget )
            {
                best = sh;
                found = true;
                break;
            }
        }
    }
    out = best;
    return found;
}

// Tile parameters for cell-sorted particles; false when the binning does not
// describe the mesh's own cells or the order has no tiled kernels (only
// orders 1 to 3 do).
bool make_tiles( TileKind kind, const MeshDev& m, int order, bool fp64,
                 const int* slot_ng, int nslot, int rec_len,
                 const cabana_amd_binning* bins, const GFieldDev& g0,
                 TileArgs& t, size_t& lds )
{
    if ( order < 1 || order > 3 || m.stag[0] || m.stag[1] || m.stag[2] )
        return false;
    if ( !bins || !bins->counts || !bins->offsets )
        return false;
    for ( int d = 0; d < 3; ++d )
        if ( bins->num_bin[d] != m.num_cell[d] )
            return false;
    const double ncell = (double)m.num_cell[0] * m.num_cell[1] * m.num_cell[2];
    const double mean_ppc = ncell > 0 ? (double)( bins->end - bins->begin ) / ncell : 0.0;
    TileShape sh;
    if ( !choose_shape( kind, fp64, slot_ng, nslot, rec_len, order, mean_ppc, sh ) )
        return false;
    const int extra = ( order == 1 ) ? 1 : 3;
    std::memset( &t, 0, sizeof( t ) );
    t.cell_offsets = bins->offsets;
    t.cell_counts = bins->counts;
    for ( int d = 0; d < 3; ++d )
    {
        t.tile[d] = sh.T[d];
        t.ntile[d] = ( m.num_cell[d] + sh.T[d] - 1 ) / sh.T[d];
        t.win[d] = sh.T[d] + extra;
    }
    t.max_runs = sh.T[0] * sh.T[1];
    // The scatter keeps a run table in one wave (<= 64 runs) and base-cell
    // coordinates in 4-bit task fields.
    if ( kind == TILE_SCATTER &&
         ( t.max_runs > 64 || sh.T[0] > 14 || sh.T[1] > 14 || sh.T[2] > 14 ) )
        return false;
    t.max_win_nodes = t.win[0] * t.win[1] * t.win[2];
    t.max_base_cells = base_cells( sh.T, order );
    t.cap = sh.cap;
    t.max_tasks = sh.max_tasks;
    t.bin_begin = bins->begin;
    // Window walk order: largest grid stride first.
    int perm[3] = { 0, 1, 2 };
    for ( int a = 0; a < 3; ++a )
        for ( int b = a + 1; b < 3; ++b )
            if ( g0.s[perm[b]] > g0.s[perm[a]] )
            {
                int tmp = perm[a];
                perm[a] = perm[b];
                perm[b] = tmp;
            }
    for ( int d = 0; d < 3; ++d )
        t.flush_perm[d] = perm[d];
    lds = tile_lds_bytes( kind, fp64, slot_ng, nslot, rec_len, order, sh );
    return true;
}

// Builds the part table of a tiling in the context scratch and sets the
// launch grid.  A binning with a build serial keeps its table in the
// context's table cache: a later call with the same binning and tile shape
// reuses it without launching anything.
cabana_amd_status make_parts( cabana_amd_context* ctx, const MeshDev& m,
                              const cabana_amd_binning* bins, TileArgs& t,
                              hipStream_t s, void* fixed = nullptr,
                              size_t fixed_bytes = 0 )
{
    using PartTable = cabana_amd_context::PartTable;
    const int64_t nbricks = (int64_t)t.ntile[0] * t.ntile[1] * t.ntile[2];
    const int64_t nbin = bins->end - bins->begin;
    const int64_t max_parts = nbricks + ( nbin + t.cap - 1 ) / t.cap;
    const int64_t grid = ( max_parts + 7 ) / 8 * 8;
    t.grid = (int32_t)( grid < max_tile_groups() ? grid : max_tile_groups() );
    PartTable* pt = nullptr;
    uint64_t gen_c = 0, gen_o = 0;
    if ( bins->serial != 0 && !fixed )
    {
        if ( bins->serial > g_serials.load() )
            return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT,
                         "binning serial was not issued by cabana_amd_new_binning_serial "
                         "(zero-initialise cabana_amd_binning)" );
        gen_c = generation_of( bins->counts );
        gen_o = generation_of( bins->offsets );
        for ( auto& e : ctx->tables )
            if ( e.serial == bins->serial && e.counts == bins->counts &&
                 e.offsets == bins->offsets && e.begin == bins->begin && e.end == bins->end &&
                 e.gen_counts == gen_c && e.gen_offsets == gen_o &&
                 e.tile[0] == t.tile[0] && e.tile[1] == t.tile[1] && e.tile[2] == t.tile[2] &&
                 e.cap == t.cap )
            {
                e.used = ++ctx->table_clock;
                t.parts = e.parts;
                t.num_parts = e.num_parts;
                return CABANA_AMD_OK;
            }
        pt = &ctx->tables[0];
        for ( auto& e : ctx->tables )
            if ( e.used < pt->used )
                pt = &e;
    }
    const int64_t nscan = ( nbricks + 1023 ) / 1024 + 1;
    size_t off = 0;
    auto carve = [&]( size_t bytes ) {
        const size_t o = off;
        off = ( off + bytes + 255 ) & ~(size_t)255;
        return o;
    };
    const size_t o_np = carve( sizeof( int32_t ) * nbricks );
    const size_t o_st = carve( sizeof( int64_t ) * nbricks );
    const size_t o_bs = carve( sizeof( int64_t ) * nscan );
    const size_t o_pt = carve( sizeof( int2 ) * max_parts );
    const size_t o_n = carve( sizeof( int32_t ) );
    char* base = nullptr;
    if ( fixed )
    {
        // The caller's fixed scratch (no allocation, no table cache).
        if ( off > fixed_bytes )
            return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT,
                         "the brick part table needs " + std::to_string( off ) +
                             " bytes; the no-allocation scratch holds " +
                             std::to_string( fixed_bytes ) );
        base = static_cast<char*>( fixed );
    }
    else if ( pt )
    {
        pt->serial = 0; // empty until rebuilt below
        if ( pt->bytes < off )
        {
            if ( pt->mem )
                hipFree( pt->mem );
            pt->mem = nullptr;
            pt->bytes = 0;
            const size_t want = off + off / 4 + 256;
            CA_HIP( ctx, hipMalloc( &pt->mem, want ), "part table" );
            pt->bytes = want;
        }
        base = static_cast<char*>( pt->mem );
    }
    else
    {
        CA_HIP( ctx, ctx->ensure( SLOT_TILES, off ), "scratch" );
        base = static_cast<char*>( ctx->scratch[SLOT_TILES] );
    }
    int2* parts = reinterpret_cast<int2*>( base + o_pt );
    int32_t* num_parts = reinterpret_cast<int32_t*>( base + o_n );
    CA_HIP( ctx,
            launch_brick_parts( *ctx, t, m.num_cell,
                                reinterpret_cast<int32_t*>( base + o_np ),
                                reinterpret_cast<int64_t*>( base + o_st ),
                                reinterpret_cast<int64_t*>( base + o_bs ), nscan, parts,
                                num_parts, s ),
            "brick parts" );
    t.parts = parts;
    t.num_parts = num_parts;
    if ( pt )
    {
        pt->counts = bins->counts;
        pt->offsets = bins->offsets;
        pt->gen_counts = gen_c;
        pt->gen_offsets = gen_o;
        pt->begin = bins->begin;
        pt->end = bins->end;
        for ( int d = 0; d < 3; ++d )
            pt->tile[d] = t.tile[d];
        pt->cap = t.cap;
        pt->parts = parts;
        pt->num_parts = num_parts;
        pt->used = ++ctx->table_clock;
        pt->serial = bins->serial;
    }
    return CABANA_AMD_OK;
}

bool bins_usable( const cabana_amd_binning* bins, int64_t n )
{
    return bins && bins->begin >= 0 && bins->begin <= bins->end &&
           bins->end <= n;
}

// Per-slot grid dofs of a pass: SHARED plans stage one 3-dof grid.
int slot_dofs( int plan, const int* forms, int nterms, int* ng )
{
    if ( plan == PLAN_MPM_G2P )
    {
        ng[0] = 3;
        return 1;
    }
    if ( plan == PLAN_MPM_ADJOINT )
    {
        // Three cotangent grids, then the shared forward grid.
        ng[0] = 1;
        ng[1] = ng[2] = ng[3] = 3;
        return 4;
    }
    for ( int i = 0; i < nterms; ++i )
        ng[i] = form_gdof( forms[i] );
    return nterms;
}

// True when a particle field is absent or stores each particle's components
// contiguously (vector_length 1): the tiled kernels then move them as vectors.
bool pf_contig( const PFieldDev& f ) { return !f.data || f.vl == 1; }

// Scatter pass over all particles: cell-register tiles over
// [bins.begin, bins.end) when the binning matches the mesh, atomics elsewhere.
cabana_amd_status scatter_pass( cabana_amd_context* ctx, int order, bool fp64,
                                int plan, ScatterArgs a, int64_t n,
                                const cabana_amd_binning* bins,
                                hipStream_t s, void* fixed = nullptr,
                                size_t fixed_bytes = 0 )
{
    int forms[MAX_TERMS], ng[MAX_TERMS];
    for ( int i = 0; i < a.nterms; ++i )
        forms[i] = a.t[i].form;
    const int nslot = slot_dofs( plan, forms, a.nterms, ng );
    a.contig = pf_contig( a.points ) && pf_contig( a.xbar );
    for ( int i = 0; i < a.nterms; ++i )
        a.contig = a.contig && pf_contig( a.t[i].pin );
    // Record header of the cell-tiled scatter: the x weights and gradients
    // with the y and z offsets for FP32 records, the three offsets for FP64.
    int rec_len = fp64 ? 3 : 2 * ( order + 1 ) + 2;
    for ( int i = 0; i < a.nterms; ++i )
        rec_len += form_pcomp( forms[i] );
    TileArgs t;
    size_t lds = 0;
    const bool tiled =
        !a.has_xbar && bins_usable( bins, n ) &&
        make_tiles( TILE_SCATTER, a.mesh, order, fp64, ng, nslot, rec_len, bins,
                    a.t[0].grid, t, lds );
    if ( !tiled )
    {
        a.begin = 0;
        a.end = n;
        CA_HIP( ctx, run_scatter_atomic( ctx, order, fp64, plan, a, s ),
                "scatter (atomic)" );
        return CABANA_AMD_OK;
    }
    t.perm = bins == &ctx->auto_bins ? ctx->auto_perm : nullptr;
    if ( bins->begin > 0 )
    {
        a.begin = 0;
        a.end = bins->begin;
        CA_HIP( ctx, run_scatter_atomic( ctx, order, fp64, plan, a, s ),
                "scatter (atomic)" );
    }
    if ( bins->end > bins->begin )
    {
        CA_CHECK( make_parts( ctx, a.mesh, bins, t, s, fixed, fixed_bytes ) );
        CA_HIP( ctx, run_scatter_cells( ctx, order, fp64, plan, a, t, lds, s ),
                "scatter (cells)" );
    }
    if ( bins->end < n )
    {
        a.begin = bins->end;
        a.end = n;
        CA_HIP( ctx, run_scatter_atomic( ctx, order, fp64, plan, a, s ),
                "scatter (atomic)" );
    }
    return CABANA_AMD_OK;
}

// Gather pass over all particles: LDS-staged tiles over [bins.begin,
// bins.end) when the binning matches the mesh, direct loads elsewhere.
cabana_amd_status gather_pass( cabana_amd_context* ctx, int order, bool fp64,
                               int plan, GatherArgs a, int64_t n,
                               const cabana_amd_binning* bins, hipStream_t s )
{
    int forms[MAX_GATHER_TERMS], ng[MAX_GATHER_TERMS];
    for ( int i = 0; i < a.nterms; ++i )
        forms[i] = a.t[i].form;
    const int nslot = slot_dofs( plan, forms, a.nterms, ng );
    a.contig = pf_contig( a.points ) && pf_contig( a.xbar );
    for ( int i = 0; i < a.nterms; ++i )
        a.contig = a.contig && pf_contig( a.t[i].pout ) && pf_contig( a.t[i].px );
    TileArgs t;
    size_t lds = 0;
    const bool tiled =
        bins_usable( bins, n ) &&
        make_tiles( TILE_GATHER, a.mesh, order, fp64, ng, nslot, 0, bins,
                    a.t[0].grid, t, lds );
    if ( !tiled )
    {
        a.begin = 0;
        a.end = n;
        CA_HIP( ctx, run_gather( ctx, order, fp64, plan, a, s ), "gather" );
        return CABANA_AMD_OK;
    }
    t.perm = bins == &ctx->auto_bins ? ctx->auto_perm : nullptr;
    if ( bins->begin > 0 )
    {
        a.begin = 0;
        a.end = bins->begin;
        CA_HIP( ctx, run_gather( ctx, order, fp64, plan, a, s ), "gather" );
    }
    if ( bins->end > bins->begin )
    {
        CA_CHECK( make_parts( ctx, a.mesh, bins, t, s ) );
        CA_HIP( ctx, run_gather_tiled( ctx, order, fp64, plan, a, t, lds, s ),
                "gather (tiled)" );
    }
    if ( bins->end < n )
    {
        a.begin = bins->end;
        a.end = n;
        CA_HIP( ctx, run_gather( ctx, order, fp64, plan, a, s ), "gather" );
    }
    return CABANA_AMD_OK;
}

// True when the first n particles of two fields share memory.  Fused gather
// passes read every output of a particle before writing any of them, so
// terms whose outputs overlap run as separate passes.
bool pfields_overlap( const PFieldDev& x, const PFieldDev& y, int64_t n, int elem )
{
    if ( !x.data || !y.data || n <= 0 )
        return false;
    auto span = [&]( const PFieldDev& f ) {
        const int64_t last = ( ( n - 1 ) / (int64_t)f.vl ) * (int64_t)f.stride +
                             ( n - 1 ) % (int64_t)f.vl + (int64_t)f.vl * ( f.ncomp - 1 );
        return ( last + 1 ) * (int64_t)elem;
    };
    const char* x0 = x.data;
    const char* y0 = y.data;
    return x0 < y0 + span( y ) && y0 < x0 + span( x );
}

cabana_amd_status validate_bins( cabana_amd_context* ctx,
                                 const cabana_amd_binning* bins, int64_t n )
{
    if ( !bins )
        return CABANA_AMD_OK;
    if ( bins->begin < 0 || bins->end < bins->begin || bins->end > n )
        return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT,
                     "binning range outside the particle range" );
    return CABANA_AMD_OK;
}

// The drop-in call shape passes no binning.  Then, for the orders with
// cell-tiled kernels, the library bins the particles itself: owned-cell
// keys, the stable radix sort of cabana_amd_linked_cell_build, counts and
// offsets, kept in SLOT_BINS with the sorted order as a permutation that
// the tiled kernels read particles through (the particles stay where they
// are).  `bins` is left as it is when the call passed one, when the
// context has the feature off, or when it cannot apply.
cabana_amd_status build_library_bins( cabana_amd_context* ctx, const MeshDev& m,
                                      const PFieldDev& pts, int64_t n, hipStream_t s );

cabana_amd_status auto_bins( cabana_amd_context* ctx, const MeshDev& m, int order,
                             const PFieldDev& pts, int64_t n,
                             const cabana_amd_binning*& bins, hipStream_t s )
{
    if ( bins || !ctx->auto_binning || order < 1 || order > 3 || n <= 0 || m.stag[0] ||
         m.stag[1] || m.stag[2] )
        return CABANA_AMD_OK;
    const int64_t ncell = (int64_t)m.num_cell[0] * m.num_cell[1] * m.num_cell[2];
    if ( ncell >= INT32_MAX / 2 )
        return CABANA_AMD_OK;
    CA_CHECK( build_library_bins( ctx, m, pts, n, s ) );
    bins = &ctx->auto_bins;
    return CABANA_AMD_OK;
}

// The library's own cell binning of particles [0, n) over the owned cells
// of a Node mesh (keys clamped into them), into ctx->auto_bins / auto_perm.
cabana_amd_status build_library_bins( cabana_amd_context* ctx, const MeshDev& m,
                                      const PFieldDev& pts, int64_t n, hipStream_t s )
{
    const int64_t ncell = (int64_t)m.num_cell[0] * m.num_cell[1] * m.num_cell[2];
    if ( ncell >= INT32_MAX / 2 )
        return fail( ctx, CABANA_AMD_ERROR_UNSUPPORTED, "more than 2^30 cells" );
    int key_bits = 0;
    while ( ( ( ncell - 1 ) >> key_bits ) != 0 )
        ++key_bits;
    int digit_bits;
    lcl_sort_passes( key_bits, &digit_bits );
    const int64_t nhist = ( (int64_t)1 << digit_bits ) * lcl_sort_tiles( n );
    const int64_t nblock = ( nhist + 1023 ) / 1024 + 1;
    const int64_t nkey = n + 64;
    CA_HIP( ctx, ctx->ensure( SLOT_CELL_OF, sizeof( int32_t ) * 4 * nkey ), "scratch" );
    CA_HIP( ctx,
            ctx->ensure( SLOT_CURSOR, ( sizeof( int64_t ) + sizeof( int32_t ) ) * ( nhist + 64 ) ),
            "scratch" );
    CA_HIP( ctx, ctx->ensure( SLOT_SCAN, sizeof( int64_t ) * nblock ), "scratch" );
    const size_t o_off = a16( sizeof( int32_t ) * (size_t)ncell );
    const size_t o_perm = a16( o_off + sizeof( int64_t ) * (size_t)ncell );
    CA_HIP( ctx, ctx->ensure( SLOT_BINS, o_perm + sizeof( int64_t ) * (size_t)n ), "scratch" );
    char* base = static_cast<char*>( ctx->scratch[SLOT_BINS] );
    int32_t* counts = reinterpret_cast<int32_t*>( base );
    int64_t* offsets = reinterpret_cast<int64_t*>( base + o_off );
    int64_t* perm = reinterpret_cast<int64_t*>( base + o_perm );
    int32_t* keys = static_cast<int32_t*>( ctx->scratch[SLOT_CELL_OF] );
    int64_t* digit_base = static_cast<int64_t*>( ctx->scratch[SLOT_CURSOR] );
    int32_t* hist = reinterpret_cast<int32_t*>( digit_base + nhist + 64 );
    int64_t* sums = static_cast<int64_t*>( ctx->scratch[SLOT_SCAN] );
    CA_HIP( ctx, launch_auto_bin_keys( *ctx, pts, n, m, keys, s ), "bin keys" );
    int32_t* sorted = keys;
    CA_HIP( ctx,
            launch_lcl_sort( *ctx, n, key_bits, keys, keys + nkey, keys + 2 * nkey,
                             keys + 3 * nkey, hist, digit_base, sums, nblock, perm, 0, &sorted,
                             s ),
            "bin sort" );
    CA_HIP( ctx, launch_lcl_bounds( *ctx, sorted, n, ncell, counts, offsets, s ), "bin bounds" );
    ctx->auto_bins = {};
    ctx->auto_bins.counts = counts;
    ctx->auto_bins.offsets = offsets;
    ctx->auto_bins.begin = 0;
    ctx->auto_bins.end = n;
    for ( int d = 0; d < 3; ++d )
        ctx->auto_bins.num_bin[d] = m.num_cell[d];
    ctx->auto_perm = perm;
    return CABANA_AMD_OK;
}

} // namespace

// ---------------------------------------------------------------------------
// Context
// ---------------------------------------------------------------------------

extern "C" int32_t cabana_amd_abi_version( void ) { return CABANA_AMD_ABI_VERSION; }

extern "C" uint64_t cabana_amd_new_binning_serial( void ) { return ++g_serials; }

extern "C" cabana_amd_status cabana_amd_context_create( int32_t device,
                                                        cabana_amd_context** out )
{
    if ( !out )
        return fail( nullptr, CABANA_AMD_ERROR_INVALID_ARGUMENT, "out is NULL" );
    *out = nullptr;
    int count = 0;
    hipError_t e = hipGetDeviceCount( &count );
    if ( e != hipSuccess || count < 1 )
        return fail( nullptr, CABANA_AMD_ERROR_DEVICE, "no HIP device available" );
    if ( device < 0 || device >= count )
        return fail( nullptr, CABANA_AMD_ERROR_INVALID_ARGUMENT,
                     "device index out of range" );
    e = hipSetDevice( device );
    if ( e != hipSuccess )
        return hip_fail( nullptr, e, "hipSetDevice" );
    auto* ctx = new cabana_amd_context();
    ctx->device = device;
    e = hipMalloc( &ctx->skipped, 2 * sizeof( unsigned long long ) );
    if ( e != hipSuccess )
    {
        delete ctx;
        return hip_fail( nullptr, e, "hipMalloc" );
    }
    ctx->bad = ctx->skipped + 1;
    e = hipMemset( ctx->skipped, 0, 2 * sizeof( unsigned long long ) );
    if ( e != hipSuccess )
    {
        hipFree( ctx->skipped );
        delete ctx;
        return hip_fail( nullptr, e, "hipMemset" );
    }
    *out = ctx;
    return CABANA_AMD_OK;
}

extern "C" cabana_amd_status cabana_amd_context_destroy( cabana_amd_context* ctx )
{
    if ( !ctx )
        return CABANA_AMD_OK;
    hipSetDevice( ctx->device );
    hipDeviceSynchronize();
    ctx->drain();
    for ( auto e : ctx->pool )
        hipEventDestroy( e );
    for ( int i = 0; i < cabana_amd_context::NUM_SLOTS; ++i )
        if ( ctx->scratch[i] )
            hipFree( ctx->scratch[i] );
    for ( auto& pt : ctx->tables )
        if ( pt.mem )
            hipFree( pt.mem );
    if ( ctx->skipped )
        hipFree( ctx->skipped );
    delete ctx;
    return CABANA_AMD_OK;
}

extern "C" const char* cabana_amd_last_error( const cabana_amd_context* ctx )
{
    if ( ctx )
        return ctx->error.c_str();
    return g_thread_error.c_str();
}

extern "C" cabana_amd_status cabana_amd_check_domain( cabana_amd_context* ctx,
                                                      cabana_amd_stream stream,
                                                      int64_t* num_skipped )
{
    if ( !ctx )
        return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT, "ctx is NULL" );
    hipStream_t s = static_cast<hipStream_t>( stream );
    CA_HIP( ctx, hipStreamSynchronize( s ), "hipStreamSynchronize" );
    unsigned long long v = 0;
    CA_HIP( ctx,
            hipMemcpy( &v, ctx->skipped, sizeof( v ), hipMemcpyDeviceToHost ),
            "hipMemcpy" );
    CA_HIP( ctx, hipMemset( ctx->skipped, 0, sizeof( v ) ), "hipMemset" );
    if ( num_skipped )
        *num_skipped = (int64_t)v;
    if ( v )
        return fail( ctx, CABANA_AMD_ERROR_OUT_OF_DOMAIN,
                     std::to_string( v ) +
                         " particles had stencils outside the ghosted grid" );
    return CABANA_AMD_OK;
}

// ---------------------------------------------------------------------------
// P2G / G2P
// ---------------------------------------------------------------------------

namespace
{
// p2g: every term added into its grid, then each distinct grid folded once.
// With `fixed`, a tiled scatter builds its part table in that memory (of
// fixed_bytes) instead of the context's allocations; nothing else the call
// uses is allocated here either (terms live on the stack).
cabana_amd_status p2g_impl( cabana_amd_context* ctx, const cabana_amd_mesh* mesh,
                            int32_t order, const cabana_amd_particle_field* points,
                            int32_t num_terms, const cabana_amd_p2g_term* terms,
                            int32_t halo_width, const cabana_amd_binning* bins,
                            cabana_amd_stream stream, void* fixed, size_t fixed_bytes )
{
    if ( !ctx )
        return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT, "ctx is NULL" );
    MeshDev m;
    CA_CHECK( make_mesh( ctx, mesh, m ) );
    CA_CHECK( check_order( ctx, order ) );
    if ( !points || num_terms < 1 || num_terms > MAX_CALL_TERMS || !terms )
        return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT,
                     "need points and 1..16 terms" );
    const int w = resolve_width( m, halo_width );
    CA_CHECK( check_width( ctx, m, w ) );
    int64_t n = points->num_particle;
    CA_CHECK( validate_bins( ctx, bins, n ) );
    bool fp64 = true;
    CA_CHECK( precision_of( ctx, terms[0].values.dtype, fp64 ) );
    const int rdt = fp64 ? CABANA_AMD_FLOAT64 : CABANA_AMD_FLOAT32;

    ScatterArgs a;
    std::memset( &a, 0, sizeof( a ) );
    ScatterTermDev T[MAX_CALL_TERMS];
    a.mesh = m;
    a.skipped = ctx->skipped;
    CA_CHECK( make_pfield( ctx, *points, 3, CABANA_AMD_FLOAT64, n, false,
                           "points", a.points ) );
    for ( int i = 0; i < num_terms; ++i )
    {
        const auto& t = terms[i];
        if ( t.op < 0 || t.op > 4 )
            return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT,
                         "unknown P2G operator" );
        const int form = p2g_form( t.op );
        T[i].form = form;
        T[i].mult = t.multiplier;
        CA_CHECK( make_pfield( ctx, t.values, form_pcomp( form ), rdt, n,
                               false, "values", T[i].pin ) );
        CA_CHECK( make_gfield( ctx, m, t.grid, form_gdof( form ), rdt, false,
                               "grid", T[i].grid ) );
    }
    for ( int i = 0; i < num_terms; ++i )
        for ( int j = 0; j < i; ++j )
            if ( T[i].grid.data == T[j].grid.data &&
                 !same_grid( T[i].grid, T[j].grid ) )
                return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT,
                             "terms share grid memory with different layouts" );
    hipStream_t s = static_cast<hipStream_t>( stream );
    if ( ctx->deterministic && !fixed )
    {
        // Deterministic: the library's own binning of every particle, then
        // one pull pass per term (each node sums its particles in a fixed
        // order and is written once); no atomics anywhere.
        if ( !node_arrays( m ) )
            return fail( ctx, CABANA_AMD_ERROR_UNSUPPORTED,
                         "deterministic P2G takes Node arrays" );
        if ( n > 0 )
        {
            CA_CHECK( build_library_bins( ctx, m, a.points, n, s ) );
            CA_HIP( ctx, run_count_outside( ctx, order, m, a.points, n, s ), "domain check" );
            PullArgs pa;
            std::memset( &pa, 0, sizeof( pa ) );
            pa.mesh = m;
            pa.points = a.points;
            pa.counts = ctx->auto_bins.counts;
            pa.offsets = ctx->auto_bins.offsets;
            pa.perm = ctx->auto_perm;
            // The MPM set (ScalarValue, VectorValue, TensorDivergence, in any
            // call order) in one pass: each candidate particle's weights are
            // formed once for the three terms.  Every other term alone.
            std::vector<bool> taken( num_terms, false );
            const int iSV = find_form( T, num_terms, FORM_SV, taken );
            const int iVV = find_form( T, num_terms, FORM_VV, taken );
            const int iTD = find_form( T, num_terms, FORM_TD, taken );
            if ( iSV >= 0 && iVV >= 0 && iTD >= 0 )
            {
                pa.nterms = 3;
                pa.term[0] = multi_term( T[iSV] );
                pa.term[1] = multi_term( T[iVV] );
                pa.term[2] = multi_term( T[iTD] );
                taken[iSV] = taken[iVV] = taken[iTD] = true;
                CA_HIP( ctx, run_pull( ctx, order, fp64, pa, s ), "p2g pull" );
            }
            for ( int i = 0; i < num_terms; ++i )
            {
                if ( taken[i] )
                    continue;
                pa.nterms = 1;
                pa.term[0] = multi_term( T[i] );
                CA_HIP( ctx, run_pull( ctx, order, fp64, pa, s ), "p2g pull" );
            }
        }
        n = 0; // nothing left for the other paths; the fold below still runs
    }
    if ( !fixed )
        CA_CHECK( auto_bins( ctx, m, order, a.points, n, bins, s ) );
    if ( n > 0 )
    {
        // The fused plans, whatever the order of the terms: (ScalarValue,
        // VectorValue, TensorDivergence) or (ScalarValue, VectorValue).
        std::vector<bool> taken( num_terms, false );
        int fused[3] = { -1, -1, -1 };
        int nfused = 0;
        const int iSV = find_form( T, num_terms, FORM_SV, taken );
        const int iVV = find_form( T, num_terms, FORM_VV, taken );
        const int iTD = find_form( T, num_terms, FORM_TD, taken );
        if ( num_terms > 1 && iSV >= 0 && iVV >= 0 )
        {
            fused[0] = iSV;
            fused[1] = iVV;
            nfused = 2;
            if ( iTD >= 0 )
                fused[nfused++] = iTD;
            for ( int k = 0; k < nfused; ++k )
                taken[fused[k]] = true;
        }
        const int rest = num_terms - nfused;
        if ( num_terms > 1 && rest > 0 && !tiled_call( bins, order, node_arrays( m ) ) )
        {
            // Untiled: every term in one pass (the spline once per particle).
            MultiTermArgs ma;
            std::memset( &ma, 0, sizeof( ma ) );
            ma.mesh = m;
            ma.points = a.points;
            ma.begin = 0;
            ma.end = n;
            ma.n_p2g = num_terms;
            for ( int i = 0; i < num_terms; ++i )
                ma.s[i] = multi_term( T[i] );
            ma.skipped = ctx->skipped;
            CA_HIP( ctx, run_multi_term( ctx, order, fp64, ma, s ), "p2g terms" );
        }
        else
        {
            if ( nfused > 0 )
            {
                ScatterArgs b = a;
                b.nterms = nfused;
                for ( int k = 0; k < nfused; ++k )
                    b.t[k] = T[fused[k]];
                CA_CHECK( scatter_pass( ctx, order, fp64,
                                        nfused == 3 ? PLAN_MPM_P2G : PLAN_MASS_MOMENTUM, b, n,
                                        bins, s, fixed, fixed_bytes ) );
            }
            for ( int i = 0; i < num_terms; ++i )
            {
                if ( taken[i] )
                    continue;
                ScatterArgs b = a;
                b.nterms = 1;
                b.t[0] = T[i];
                CA_CHECK( scatter_pass( ctx, order, fp64, PLAN_SINGLE, b, n,
                                        bins, s, fixed, fixed_bytes ) );
            }
        }
    }
    // Fold every distinct grid once.
    for ( int i = 0; i < num_terms; ++i )
    {
        bool seen = false;
        for ( int j = 0; j < i; ++j )
            seen = seen || T[j].grid.data == T[i].grid.data;
        if ( !seen )
            CA_HIP( ctx,
                    launch_halo_scatter_sum( *ctx, fp64, m, T[i].grid, w, s ),
                    "halo scatter" );
    }
    return CABANA_AMD_OK;
}
} // namespace

extern "C" cabana_amd_status cabana_amd_p2g(
    cabana_amd_context* ctx, const cabana_amd_mesh* mesh, int32_t order,
    const cabana_amd_particle_field* points, int32_t num_terms,
    const cabana_amd_p2g_term* terms, int32_t halo_width,
    const cabana_amd_binning* bins, cabana_amd_stream stream )
{
    TraceRange trace_range( "Cabana::Grid::p2g" );
    return p2g_impl( ctx, mesh, order, points, num_terms, terms, halo_width, bins, stream,
                     nullptr, 0 );
}

namespace cabana_amd
{
cabana_amd_status p2g_fixed_scratch( cabana_amd_context* ctx, const cabana_amd_mesh* mesh,
                                     int32_t order, const cabana_amd_particle_field* points,
                                     int32_t num_terms, const cabana_amd_p2g_term* terms,
                                     int32_t halo_width, const cabana_amd_binning* bins,
                                     cabana_amd_stream stream, void* scratch,
                                     size_t scratch_bytes )
{
    if ( !ctx )
        return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT, "ctx is NULL" );
    int cur = -1;
    if ( hipGetDevice( &cur ) != hipSuccess || cur != ctx->device )
        return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT,
                     "the context's device must be the current device" );
    if ( !scratch )
        return fail( ctx, CABANA_AMD_ERROR_DEVICE,
                     "the fixed scratch of this device has no address" );
    // The kernel timer creates events: it stays off for this call.
    const bool prof = ctx->profiling;
    ctx->profiling = false;
    const cabana_amd_status st = p2g_impl( ctx, mesh, order, points, num_terms, terms,
                                           halo_width, bins, stream, scratch, scratch_bytes );
    ctx->profiling = prof;
    return st;
}

cabana_amd_status fail_with( cabana_amd_context* ctx, cabana_amd_status st, const char* msg )
{
    return fail( ctx, st, msg );
}
} // namespace cabana_amd

namespace
{
// The passes of a G2P call.  (VectorValue, VectorGradient) from one grid
// into distinct fields is the fused plan, in either order; without tiles
// any other set runs in one pass; with tiles the other terms run one pass
// each.
cabana_amd_status g2p_passes( cabana_amd_context* ctx, int order, bool fp64, const GatherArgs& a,
                              const GatherTermDev* T, int num_terms, int64_t n,
                              const cabana_amd_binning* bins, hipStream_t s )
{
    std::vector<bool> taken( num_terms, false );
    int pair[2] = { -1, -1 };
    for ( int i = 0; i < num_terms && pair[0] < 0; ++i )
        for ( int j = 0; j < num_terms; ++j )
            if ( T[i].form == FORM_VV && T[j].form == FORM_TDT && same_grid( T[i].grid, T[j].grid ) &&
                 !pfields_overlap( T[i].pout, T[j].pout, n, fp64 ? 8 : 4 ) )
            {
                pair[0] = i;
                pair[1] = j;
                break;
            }
    const int nfused = pair[0] >= 0 ? 2 : 0;
    if ( nfused )
        taken[pair[0]] = taken[pair[1]] = true;
    if ( num_terms > 1 && num_terms > nfused && !tiled_call( bins, order, node_arrays( a.mesh ) ) )
    {
        MultiTermArgs ma;
        std::memset( &ma, 0, sizeof( ma ) );
        ma.mesh = a.mesh;
        ma.points = a.points;
        ma.begin = 0;
        ma.end = n;
        ma.n_g2p = num_terms;
        for ( int i = 0; i < num_terms; ++i )
            ma.g[i] = multi_term( T[i] );
        ma.skipped = ctx->skipped;
        CA_HIP( ctx, run_multi_term( ctx, order, fp64, ma, s ), "g2p terms" );
        return CABANA_AMD_OK;
    }
    if ( num_terms == 1 || nfused )
    {
        GatherArgs b = a;
        b.nterms = num_terms == 1 ? 1 : 2;
        b.t[0] = T[num_terms == 1 ? 0 : pair[0]];
        if ( nfused )
            b.t[1] = T[pair[1]];
        CA_CHECK( gather_pass( ctx, order, fp64, nfused ? PLAN_MPM_G2P : PLAN_SINGLE, b, n, bins,
                               s ) );
    }
    for ( int i = 0; i < num_terms && num_terms > 1; ++i )
    {
        if ( taken[i] )
            continue;
        GatherArgs b = a;
        b.nterms = 1;
        b.t[0] = T[i];
        CA_CHECK( gather_pass( ctx, order, fp64, PLAN_SINGLE, b, n, bins, s ) );
    }
    return CABANA_AMD_OK;
}
} // namespace

extern "C" cabana_amd_status cabana_amd_g2p(
    cabana_amd_context* ctx, const cabana_amd_mesh* mesh, int32_t order,
    const cabana_amd_particle_field* points, int32_t num_terms,
    const cabana_amd_g2p_term* terms, int32_t halo_width,
    const cabana_amd_binning* bins, cabana_amd_stream stream )
{
    TraceRange trace_range( "Cabana::Grid::g2p" );
    if ( !ctx )
        return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT, "ctx is NULL" );
    MeshDev m;
    CA_CHECK( make_mesh( ctx, mesh, m ) );
    CA_CHECK( check_order( ctx, order ) );
    if ( !points || num_terms < 1 || num_terms > MAX_CALL_TERMS || !terms )
        return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT,
                     "need points and 1..16 terms" );
    const int w = resolve_width( m, halo_width );
    CA_CHECK( check_width( ctx, m, w ) );
    const int64_t n = points->num_particle;
    CA_CHECK( validate_bins( ctx, bins, n ) );
    bool fp64 = true;
    CA_CHECK( precision_of( ctx, terms[0].values.dtype, fp64 ) );
    const int rdt = fp64 ? CABANA_AMD_FLOAT64 : CABANA_AMD_FLOAT32;

    GatherArgs a;
    std::memset( &a, 0, sizeof( a ) );
    std::vector<GatherTermDev> T( num_terms );
    a.mesh = m;
    a.skipped = ctx->skipped;
    CA_CHECK( make_pfield( ctx, *points, 3, CABANA_AMD_FLOAT64, n, false,
                           "points", a.points ) );
    for ( int i = 0; i < num_terms; ++i )
    {
        const auto& t = terms[i];
        if ( t.op < 0 || t.op > 4 )
            return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT,
                         "unknown G2P operator" );
        const int form = g2p_form( t.op );
        T[i].form = form;
        T[i].mult = t.multiplier;
        CA_CHECK( make_pfield( ctx, t.values, form_pcomp( form ), rdt, n,
                               false, "values", T[i].pout ) );
        CA_CHECK( make_gfield( ctx, m, t.grid, form_gdof( form ), rdt, false,
                               "grid", T[i].grid ) );
    }
    hipStream_t s = static_cast<hipStream_t>( stream );
    CA_CHECK( auto_bins( ctx, m, order, a.points, n, bins, s ) );
    for ( int i = 0; i < num_terms; ++i )
    {
        bool seen = false;
        for ( int j = 0; j < i; ++j )
            seen = seen || T[j].grid.data == T[i].grid.data;
        if ( !seen )
            CA_HIP( ctx, launch_halo_gather( *ctx, fp64, m, T[i].grid, w, s ),
                    "halo gather" );
    }
    if ( n <= 0 )
        return CABANA_AMD_OK;
    return g2p_passes( ctx, order, fp64, a, T.data(), num_terms, n, bins, s );
}

// ---------------------------------------------------------------------------
// G2P followed by P2G in one call
// ---------------------------------------------------------------------------

namespace
{
// Bytes from a grid's first element to one past its last.
size_t grid_span( const GFieldDev& g, bool fp64 )
{
    int64_t last = 0;
    for ( int d = 0; d < 4; ++d )
        last += (int64_t)( g.ext[d] - 1 ) * g.s[d];
    return (size_t)( last + 1 ) * ( fp64 ? 8 : 4 );
}

bool grids_overlap( const GFieldDev& x, const GFieldDev& y, bool fp64 )
{
    return x.data < y.data + grid_span( y, fp64 ) && y.data < x.data + grid_span( x, fp64 );
}
} // namespace

extern "C" cabana_amd_status cabana_amd_g2p2g(
    cabana_amd_context* ctx, const cabana_amd_mesh* mesh, int32_t order,
    const cabana_amd_particle_field* points, int32_t num_g2p_terms,
    const cabana_amd_g2p_term* g2p_terms, int32_t num_p2g_terms,
    const cabana_amd_p2g_term* p2g_terms, int32_t halo_width,
    const cabana_amd_binning* bins, cabana_amd_stream stream )
{
    TraceRange trace_range( "Cabana::Grid::g2p2g" );
    if ( !ctx )
        return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT, "ctx is NULL" );
    if ( !points || num_g2p_terms < 1 || num_g2p_terms > MAX_CALL_TERMS || !g2p_terms ||
         num_p2g_terms < 1 || num_p2g_terms > MAX_CALL_TERMS || !p2g_terms )
        return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT,
                     "need points, 1..16 G2P terms and 1..16 P2G terms" );
    if ( ctx->deterministic )
    {
        // Every G2P read before any P2G write; the P2G is the deterministic one.
        CA_CHECK( cabana_amd_g2p( ctx, mesh, order, points, num_g2p_terms, g2p_terms, halo_width,
                                  bins, stream ) );
        return cabana_amd_p2g( ctx, mesh, order, points, num_p2g_terms, p2g_terms, halo_width,
                               bins, stream );
    }
    if ( !bins && ctx->auto_binning && order >= 1 && order <= 3 )
    {
        // The library's own binning, built once for both halves.
        MeshDev mb;
        CA_CHECK( make_mesh( ctx, mesh, mb ) );
        PFieldDev pb;
        CA_CHECK( make_pfield( ctx, *points, 3, CABANA_AMD_FLOAT64, points->num_particle, false,
                               "points", pb ) );
        CA_CHECK( auto_bins( ctx, mb, order, pb, points->num_particle, bins,
                             static_cast<hipStream_t>( stream ) ) );
    }
    if ( tiled_call( bins, order, mesh && mesh->entity == CABANA_AMD_NODE ) )
    {
        // Cell-tiled: the G2P passes, then the P2G passes.  Every read of a
        // grid precedes every addition into it, as in the reference's
        // g2p2g with a duplicated (Serial, OpenMP) scatter view.
        CA_CHECK( cabana_amd_g2p( ctx, mesh, order, points, num_g2p_terms, g2p_terms,
                                  halo_width, bins, stream ) );
        return cabana_amd_p2g( ctx, mesh, order, points, num_p2g_terms, p2g_terms, halo_width,
                               bins, stream );
    }
    MeshDev m;
    CA_CHECK( make_mesh( ctx, mesh, m ) );
    CA_CHECK( check_order( ctx, order ) );
    const int w = resolve_width( m, halo_width );
    CA_CHECK( check_width( ctx, m, w ) );
    const int64_t n = points->num_particle;
    CA_CHECK( validate_bins( ctx, bins, n ) );
    bool fp64 = true;
    CA_CHECK( precision_of( ctx, g2p_terms[0].values.dtype, fp64 ) );
    const int rdt = fp64 ? CABANA_AMD_FLOAT64 : CABANA_AMD_FLOAT32;
    MultiTermArgs a;
    std::memset( &a, 0, sizeof( a ) );
    a.mesh = m;
    a.skipped = ctx->skipped;
    a.begin = 0;
    a.end = n;
    a.n_g2p = num_g2p_terms;
    a.n_p2g = num_p2g_terms;
    CA_CHECK( make_pfield( ctx, *points, 3, CABANA_AMD_FLOAT64, n, false, "points", a.points ) );
    for ( int i = 0; i < num_g2p_terms; ++i )
    {
        const auto& t = g2p_terms[i];
        if ( t.op < 0 || t.op > 4 )
            return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT, "unknown G2P operator" );
        const int form = g2p_form( t.op );
        a.g[i].form = form;
        a.g[i].mult = t.multiplier;
        CA_CHECK( make_pfield( ctx, t.values, form_pcomp( form ), rdt, n, false, "values",
                               a.g[i].f ) );
        CA_CHECK( make_gfield( ctx, m, t.grid, form_gdof( form ), rdt, false, "grid",
                               a.g[i].grid ) );
    }
    for ( int i = 0; i < num_p2g_terms; ++i )
    {
        const auto& t = p2g_terms[i];
        if ( t.op < 0 || t.op > 4 )
            return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT, "unknown P2G operator" );
        const int form = p2g_form( t.op );
        a.s[i].form = form;
        a.s[i].mult = t.multiplier;
        CA_CHECK( make_pfield( ctx, t.values, form_pcomp( form ), rdt, n, false, "values",
                               a.s[i].f ) );
        CA_CHECK( make_gfield( ctx, m, t.grid, form_gdof( form ), rdt, false, "grid",
                               a.s[i].grid ) );
    }
    hipStream_t s = static_cast<hipStream_t>( stream );
    // Halo::gather of every distinct G2P grid (Cabana_Grid_Interpolation
    // Performance.cpp:140-141 gathers before the point loop).
    for ( int i = 0; i < num_g2p_terms; ++i )
    {
        bool seen = false;
        for ( int j = 0; j < i; ++j )
            seen = seen || a.g[j].grid.data == a.g[i].grid.data;
        if ( !seen )
            CA_HIP( ctx, launch_halo_gather( *ctx, fp64, m, a.g[i].grid, w, s ), "halo gather" );
    }
    if ( n > 0 )
    {
        // A G2P grid that a P2G term also adds into is read from a copy
        // taken after the gather, so that every particle reads the values
        // the call started from.
        size_t need = 0;
        std::vector<int> copied( num_g2p_terms, -1 );
        for ( int i = 0; i < num_g2p_terms; ++i )
        {
            bool hit = false;
            for ( int j = 0; j < num_p2g_terms; ++j )
                hit = hit || grids_overlap( a.g[i].grid, a.s[j].grid, fp64 );
            if ( !hit )
                continue;
            for ( int j = 0; j < i; ++j )
                if ( copied[j] >= 0 && a.g[j].grid.data == a.g[i].grid.data )
                    copied[i] = copied[j];
            if ( copied[i] < 0 )
            {
                copied[i] = (int)need;
                need = a16( need + grid_span( a.g[i].grid, fp64 ) );
            }
        }
        if ( need > 0 )
        {
            CA_HIP( ctx, ctx->ensure( SLOT_SNAPSHOT, need ), "scratch" );
            char* base = static_cast<char*>( ctx->scratch[SLOT_SNAPSHOT] );
            for ( int i = 0; i < num_g2p_terms; ++i )
            {
                if ( copied[i] < 0 )
                    continue;
                bool first = true;
                for ( int j = 0; j < i; ++j )
                    first = first && !( copied[j] == copied[i] );
                if ( first )
                    CA_HIP( ctx,
                            hipMemcpyAsync( base + copied[i], a.g[i].grid.data,
                                            grid_span( a.g[i].grid, fp64 ),
                                            hipMemcpyDeviceToDevice, s ),
                            "hipMemcpyAsync" );
                a.g[i].grid.data = base + copied[i];
            }
        }
        CA_HIP( ctx, run_multi_term( ctx, order, fp64, a, s ), "g2p2g" );
    }
    // Halo::scatter(Sum) of every distinct P2G grid.
    for ( int i = 0; i < num_p2g_terms; ++i )
    {
        bool seen = false;
        for ( int j = 0; j < i; ++j )
            seen = seen || a.s[j].grid.data == a.s[i].grid.data;
        if ( !seen )
            CA_HIP( ctx, launch_halo_scatter_sum( *ctx, fp64, m, a.s[i].grid, w, s ),
                    "halo scatter" );
    }
    return CABANA_AMD_OK;
}

// ---------------------------------------------------------------------------
// Adjoints
// ---------------------------------------------------------------------------

namespace
{
// cabana_amd_p2g_adjoint, and with num_gx > 0 the position part of the G2P
// adjoint terms gx as well.  Those are taken into the pass (and *fused set)
// only for the MPM set: P2G (ScalarValue, VectorValue, TensorDivergence)
// with G2P (VectorValue, VectorGradient) on one grid that is none of the
// cotangent grids; otherwise *fused stays false and the caller runs them.
cabana_amd_status p2g_adjoint_impl( cabana_amd_context* ctx, const cabana_amd_mesh* mesh,
                                    int32_t order, const cabana_amd_particle_field* points,
                                    int32_t num_terms, const cabana_amd_p2g_adjoint_term* terms,
                                    int32_t halo_width,
                                    const cabana_amd_particle_field* points_bar,
                                    const cabana_amd_binning* bins, cabana_amd_stream stream,
                                    int32_t num_gx, const cabana_amd_g2p_adjoint_term* gx,
                                    bool* fused )
{
    if ( fused )
        *fused = false;
    if ( !ctx )
        return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT, "ctx is NULL" );
    MeshDev m;
    CA_CHECK( make_mesh( ctx, mesh, m ) );
    CA_CHECK( check_order( ctx, order ) );
    if ( !points || num_terms < 1 || num_terms > MAX_CALL_TERMS || !terms )
        return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT,
                     "need points and 1..16 terms" );
    const int w = resolve_width( m, halo_width );
    CA_CHECK( check_width( ctx, m, w ) );
    const int64_t n = points->num_particle;
    CA_CHECK( validate_bins( ctx, bins, n ) );
    bool fp64 = true;
    CA_CHECK( precision_of( ctx, terms[0].grid_bar.dtype, fp64 ) );
    const int rdt = fp64 ? CABANA_AMD_FLOAT64 : CABANA_AMD_FLOAT32;
    const bool xbar = points_bar && points_bar->data;

    GatherArgs a;
    std::memset( &a, 0, sizeof( a ) );
    std::vector<GatherTermDev> T( num_terms );
    a.mesh = m;
    a.skipped = ctx->skipped;
    CA_CHECK( make_pfield( ctx, *points, 3, CABANA_AMD_FLOAT64, n, false,
                           "points", a.points ) );
    if ( xbar )
        CA_CHECK( make_pfield( ctx, *points_bar, 3, CABANA_AMD_FLOAT64, n,
                               false, "points_bar", a.xbar ) );
    a.has_xbar = xbar ? 1 : 0;
    for ( int i = 0; i < num_terms; ++i )
    {
        const auto& t = terms[i];
        if ( t.op < 0 || t.op > 4 )
            return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT,
                         "unknown P2G operator" );
        const int form = p2g_form( t.op );
        T[i].form = form;
        T[i].mult = t.multiplier;
        CA_CHECK( make_pfield( ctx, t.values_bar, form_pcomp( form ), rdt, n,
                               true, "values_bar", T[i].pout ) );
        CA_CHECK( make_pfield( ctx, t.values, form_pcomp( form ), rdt, n,
                               !xbar, "values", T[i].px ) );
        CA_CHECK( make_gfield( ctx, m, t.grid_bar, form_gdof( form ), rdt,
                               false, "grid_bar", T[i].grid ) );
    }
    hipStream_t s = static_cast<hipStream_t>( stream );
    CA_CHECK( auto_bins( ctx, m, order, a.points, n, bins, s ) );
    for ( int i = 0; i < num_terms; ++i )
    {
        bool seen = false;
        for ( int j = 0; j < i; ++j )
            seen = seen || T[j].grid.data == T[i].grid.data;
        if ( !seen )
            CA_HIP( ctx,
                    launch_halo_adjoint_fill( *ctx, fp64, m, T[i].grid, w, s ),
                    "halo adjoint fill" );
    }
    if ( n <= 0 )
        return CABANA_AMD_OK;
    const bool mpm = num_terms == 3 && xbar && T[0].form == FORM_SV &&
                     T[1].form == FORM_VV && T[2].form == FORM_TD && T[0].pout.data &&
                     T[1].pout.data && T[2].pout.data &&
                     !pfields_overlap( T[0].pout, T[1].pout, n, fp64 ? 8 : 4 ) &&
                     !pfields_overlap( T[0].pout, T[2].pout, n, fp64 ? 8 : 4 ) &&
                     !pfields_overlap( T[1].pout, T[2].pout, n, fp64 ? 8 : 4 );
    if ( mpm && num_gx == 2 && gx && fused && gx[0].op == CABANA_AMD_VECTOR_VALUE_G2P &&
         gx[1].op == CABANA_AMD_VECTOR_GRADIENT_G2P )
    {
        // The position part of the G2P adjoint of (VectorValue,
        // VectorGradient) as two position-only terms of the same gather:
        // the output cotangents contracted with the forward grid's jets.
        GatherTermDev X[2];
        std::memset( X, 0, sizeof( X ) );
        for ( int i = 0; i < 2; ++i )
        {
            const int form = g2p_form( gx[i].op );
            X[i].form = form;
            X[i].mult = gx[i].multiplier;
            CA_CHECK( make_pfield( ctx, gx[i].values_bar, form_pcomp( form ), rdt, n, false,
                                   "values_bar", X[i].px ) );
            CA_CHECK( make_gfield( ctx, m, gx[i].grid, form_gdof( form ), rdt, false, "grid",
                                   X[i].grid ) );
        }
        bool own_grid = same_grid( X[0].grid, X[1].grid );
        for ( int i = 0; i < num_terms; ++i )
            own_grid = own_grid && T[i].grid.data != X[0].grid.data;
        if ( own_grid )
        {
            // Forward ghost fill of the grid read for the position cotangent.
            CA_HIP( ctx, launch_halo_gather( *ctx, fp64, m, X[0].grid, w, s ), "halo gather" );
            a.nterms = 5;
            for ( int i = 0; i < 3; ++i )
                a.t[i] = T[i];
            a.t[3] = X[0];
            a.t[4] = X[1];
            *fused = true;
            return gather_pass( ctx, order, fp64, PLAN_MPM_ADJOINT, a, n, bins, s );
        }
    }
    if ( num_terms == 1 || mpm )
    {
        a.nterms = num_terms;
        for ( int i = 0; i < num_terms; ++i )
            a.t[i] = T[i];
        return gather_pass( ctx, order, fp64, mpm ? PLAN_MPM_P2G : PLAN_SINGLE,
                            a, n, bins, s );
    }
    for ( int i = 0; i < num_terms; ++i )
    {
        GatherArgs b = a;
        b.nterms = 1;
        b.t[0] = T[i];
        CA_CHECK( gather_pass( ctx, order, fp64, PLAN_SINGLE, b, n, bins, s ) );
    }
    return CABANA_AMD_OK;
}

} // namespace

extern "C" cabana_amd_status cabana_amd_p2g_adjoint(
    cabana_amd_context* ctx, const cabana_amd_mesh* mesh, int32_t order,
    const cabana_amd_particle_field* points, int32_t num_terms,
    const cabana_amd_p2g_adjoint_term* terms, int32_t halo_width,
    const cabana_amd_particle_field* points_bar,
    const cabana_amd_binning* bins, cabana_amd_stream stream )
{
    TraceRange trace_range( "cabana_amd::p2g_adjoint" );
    return p2g_adjoint_impl( ctx, mesh, order, points, num_terms, terms, halo_width, points_bar,
                             bins, stream, 0, nullptr, nullptr );
}

extern "C" cabana_amd_status cabana_amd_adjoint_gather(
    cabana_amd_context* ctx, const cabana_amd_mesh* mesh, int32_t order,
    const cabana_amd_particle_field* points, int32_t num_p2g_terms,
    const cabana_amd_p2g_adjoint_term* p2g_terms, int32_t num_g2p_terms,
    const cabana_amd_g2p_adjoint_term* g2p_terms, int32_t halo_width,
    const cabana_amd_particle_field* points_bar, const cabana_amd_binning* bins,
    cabana_amd_stream stream )
{
    TraceRange trace_range( "cabana_amd::adjoint_gather" );
    if ( !ctx )
        return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT, "ctx is NULL" );
    if ( !points || !points_bar || !points_bar->data || num_p2g_terms < 1 ||
         num_p2g_terms > MAX_CALL_TERMS || !p2g_terms || num_g2p_terms < 1 ||
         num_g2p_terms > MAX_CALL_TERMS || !g2p_terms )
        return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT,
                     "need points, points_bar, 1..16 P2G terms and 1..16 G2P terms" );
    bool fused = false;
    CA_CHECK( p2g_adjoint_impl( ctx, mesh, order, points, num_p2g_terms, p2g_terms, halo_width,
                                points_bar, bins, stream, num_g2p_terms, g2p_terms, &fused ) );
    if ( fused )
        return CABANA_AMD_OK;
    // Otherwise the G2P adjoint's position part as its own passes: the
    // terms without their grid cotangents.
    cabana_amd_g2p_adjoint_term gt[MAX_CALL_TERMS];
    for ( int i = 0; i < num_g2p_terms; ++i )
    {
        gt[i] = g2p_terms[i];
        std::memset( &gt[i].grid_bar, 0, sizeof( gt[i].grid_bar ) );
    }
    return cabana_amd_g2p_adjoint( ctx, mesh, order, points, num_g2p_terms, gt, halo_width,
                                   points_bar, bins, stream );
}

extern "C" cabana_amd_status cabana_amd_g2p_adjoint(
    cabana_amd_context* ctx, const cabana_amd_mesh* mesh, int32_t order,
    const cabana_amd_particle_field* points, int32_t num_terms,
    const cabana_amd_g2p_adjoint_term* terms, int32_t halo_width,
    const cabana_amd_particle_field* points_bar,
    const cabana_amd_binning* bins, cabana_amd_stream stream )
{
    TraceRange trace_range( "cabana_amd::g2p_adjoint" );
    if ( !ctx )
        return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT, "ctx is NULL" );
    MeshDev m;
    CA_CHECK( make_mesh( ctx, mesh, m ) );
    CA_CHECK( check_order( ctx, order ) );
    if ( !points || num_terms < 1 || num_terms > MAX_CALL_TERMS || !terms )
        return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT,
                     "need points and 1..16 terms" );
    const int w = resolve_width( m, halo_width );
    CA_CHECK( check_width( ctx, m, w ) );
    const int64_t n = points->num_particle;
    CA_CHECK( validate_bins( ctx, bins, n ) );
    bool fp64 = true;
    CA_CHECK( precision_of( ctx, terms[0].grid.dtype, fp64 ) );
    const int rdt = fp64 ? CABANA_AMD_FLOAT64 : CABANA_AMD_FLOAT32;
    const bool xbar = points_bar && points_bar->data;

    ScatterArgs a;
    std::memset( &a, 0, sizeof( a ) );
    std::vector<ScatterTermDev> T( num_terms );
    a.mesh = m;
    a.skipped = ctx->skipped;
    CA_CHECK( make_pfield( ctx, *points, 3, CABANA_AMD_FLOAT64, n, false,
                           "points", a.points ) );
    if ( xbar )
        CA_CHECK( make_pfield( ctx, *points_bar, 3, CABANA_AMD_FLOAT64, n,
                               false, "points_bar", a.xbar ) );
    a.has_xbar = xbar ? 1 : 0;
    for ( int i = 0; i < num_terms; ++i )
    {
        const auto& t = terms[i];
        if ( t.op < 0 || t.op > 4 )
            return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT,
                         "unknown G2P operator" );
        const int form = g2p_form( t.op );
        T[i].form = form;
        T[i].mult = t.multiplier;
        CA_CHECK( make_pfield( ctx, t.values_bar, form_pcomp( form ), rdt, n,
                               false, "values_bar", T[i].pin ) );
        CA_CHECK( make_gfield( ctx, m, t.grid_bar, form_gdof( form ), rdt,
                               true, "grid_bar", T[i].grid ) );
        CA_CHECK( make_gfield( ctx, m, t.grid, form_gdof( form ), rdt, !xbar,
                               "grid", T[i].gx ) );
        if ( !T[i].grid.data && !xbar )
            return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT,
                         "term has neither grid_bar nor points_bar" );
    }
    hipStream_t s = static_cast<hipStream_t>( stream );
    CA_CHECK( auto_bins( ctx, m, order, a.points, n, bins, s ) );
    // Forward ghost fill of the grids read for the position cotangent.
    if ( xbar )
        for ( int i = 0; i < num_terms; ++i )
        {
            bool seen = false;
            for ( int j = 0; j < i; ++j )
                seen = seen || T[j].gx.data == T[i].gx.data;
            if ( !seen )
                CA_HIP( ctx,
                        launch_halo_gather( *ctx, fp64, m, T[i].gx, w, s ),
                        "halo gather" );
        }
    // Cotangent grids: clear ghosts, scatter, fold.
    for ( int i = 0; i < num_terms; ++i )
    {
        bool seen = false;
        for ( int j = 0; j < i; ++j )
            seen = seen || T[j].grid.data == T[i].grid.data;
        if ( !seen && T[i].grid.data )
            CA_HIP( ctx, launch_zero_ghosts( *ctx, fp64, m, T[i].grid, s ),
                    "zero ghosts" );
    }
    if ( n > 0 && xbar )
    {
        // Position cotangents: a G2P-shaped gather of the forward grid,
        // contracted with the output cotangents (no value output).
        GatherArgs ga;
        std::memset( &ga, 0, sizeof( ga ) );
        ga.mesh = m;
        ga.skipped = ctx->skipped;
        ga.points = a.points;
        ga.xbar = a.xbar;
        ga.has_xbar = 1;
        std::vector<GatherTermDev> GT( num_terms );
        for ( int i = 0; i < num_terms; ++i )
        {
            std::memset( &GT[i], 0, sizeof( GT[i] ) );
            GT[i].form = T[i].form;
            GT[i].mult = T[i].mult;
            GT[i].px = T[i].pin;
            GT[i].grid = T[i].gx;
        }
        const bool gpair = num_terms == 2 && T[0].form == FORM_VV &&
                           T[1].form == FORM_TDT &&
                           same_grid( T[0].gx, T[1].gx );
        if ( num_terms == 1 || gpair )
        {
            ga.nterms = num_terms;
            for ( int i = 0; i < num_terms; ++i )
                ga.t[i] = GT[i];
            CA_CHECK( gather_pass( ctx, order, fp64,
                                   gpair ? PLAN_MPM_G2P : PLAN_SINGLE, ga, n,
                                   bins, s ) );
        }
        else
        {
            for ( int i = 0; i < num_terms; ++i )
            {
                GatherArgs b = ga;
                b.nterms = 1;
                b.t[0] = GT[i];
                CA_CHECK( gather_pass( ctx, order, fp64, PLAN_SINGLE, b, n,
                                       bins, s ) );
            }
        }
    }
    // Grid cotangents: a P2G-shaped scatter of the output cotangents.
    a.has_xbar = 0;
    if ( n > 0 )
    {
        const bool pair = num_terms == 2 && T[0].form == FORM_VV &&
                          T[1].form == FORM_TDT && T[0].grid.data &&
                          same_grid( T[0].grid, T[1].grid );
        if ( num_terms == 1 || pair )
        {
            a.nterms = num_terms;
            for ( int i = 0; i < num_terms; ++i )
                a.t[i] = T[i];
            if ( T[0].grid.data )
                CA_CHECK( scatter_pass( ctx, order, fp64,
                                        pair ? PLAN_MPM_G2P : PLAN_SINGLE, a, n,
                                        bins, s ) );
        }
        else
        {
            for ( int i = 0; i < num_terms; ++i )
            {
                if ( !T[i].grid.data )
                    continue;
                ScatterArgs b = a;
                b.nterms = 1;
                b.t[0] = T[i];
                CA_CHECK( scatter_pass( ctx, order, fp64, PLAN_SINGLE, b, n,
                                        bins, s ) );
            }
        }
    }
    for ( int i = 0; i < num_terms; ++i )
    {
        bool seen = false;
        for ( int j = 0; j < i; ++j )
            seen = seen || T[j].grid.data == T[i].grid.data;
        if ( !seen && T[i].grid.data )
            CA_HIP( ctx,
                    launch_halo_scatter_sum( *ctx, fp64, m, T[i].grid, w, s ),
                    "halo scatter" );
    }
    return CABANA_AMD_OK;
}

// ---------------------------------------------------------------------------
// Halo
// ---------------------------------------------------------------------------

namespace
{
cabana_amd_status halo_common( cabana_amd_context* ctx,
                               const cabana_amd_mesh* mesh,
                               const cabana_amd_grid_field* grid,
                               int32_t halo_width, MeshDev& m, GFieldDev& g,
                               int& w, bool& fp64 )
{
    if ( !ctx )
        return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT, "ctx is NULL" );
    CA_CHECK( make_mesh( ctx, mesh, m ) );
    if ( !grid )
        return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT, "grid is NULL" );
    CA_CHECK( precision_of( ctx, grid->dtype, fp64 ) );
    CA_CHECK( make_gfield( ctx, m, *grid, -1, grid->dtype, false, "grid", g ) );
    w = resolve_width( m, halo_width );
    return check_width( ctx, m, w );
}
} // namespace

extern "C" cabana_amd_status cabana_amd_halo_gather(
    cabana_amd_context* ctx, const cabana_amd_mesh* mesh,
    const cabana_amd_grid_field* grid, int32_t halo_width,
    cabana_amd_stream stream )
{
    TraceRange trace_range( "Cabana::Grid::gather" );
    MeshDev m;
    GFieldDev g;
    int w;
    bool fp64;
    CA_CHECK( halo_common( ctx, mesh, grid, halo_width, m, g, w, fp64 ) );
    CA_HIP( ctx,
            launch_halo_gather( *ctx, fp64, m, g, w,
                                static_cast<hipStream_t>( stream ) ),
            "halo gather" );
    return CABANA_AMD_OK;
}

extern "C" cabana_amd_status cabana_amd_halo_scatter_sum(
    cabana_amd_context* ctx, const cabana_amd_mesh* mesh,
    const cabana_amd_grid_field* grid, int32_t halo_width,
    cabana_amd_stream stream )
{
    TraceRange trace_range( "Cabana::Grid::scatter" );
    MeshDev m;
    GFieldDev g;
    int w;
    bool fp64;
    CA_CHECK( halo_common( ctx, mesh, grid, halo_width, m, g, w, fp64 ) );
    CA_HIP( ctx,
            launch_halo_scatter_sum( *ctx, fp64, m, g, w,
                                     static_cast<hipStream_t>( stream ) ),
            "halo scatter" );
    return CABANA_AMD_OK;
}

// ---------------------------------------------------------------------------
// Binning
// ---------------------------------------------------------------------------

namespace
{
// CartesianGrid::init (core/src/impl/Cabana_CartesianGrid.hpp:61-73,279-283).
bool cart_init( const double delta[3], const double min[3],
                const double max[3], CartGridDev& g, int64_t& ncell )
{
    ncell = 1;
    for ( int d = 0; d < 3; ++d )
    {
        if ( !( delta[d] > 0.0 ) || !( max[d] > min[d] ) )
            return false;
        const double nx = std::floor( ( max[d] - min[d] ) * ( 1.0 / delta[d] ) );
        if ( nx < 1.0 || nx > 2.0e9 )
            return false;
        g.nx[d] = (int32_t)nx;
        const double dx = ( max[d] - min[d] ) / g.nx[d];
        g.rdx[d] = 1.0 / dx;
        g.min[d] = min[d];
        ncell *= g.nx[d];
    }
    return ncell <= ( int64_t( 1 ) << 31 ) - 1;
}
} // namespace

extern "C" cabana_amd_status cabana_amd_linked_cell_num_bins(
    const double grid_delta[3], const double grid_min[3],
    const double grid_max[3], int32_t num_bin[3] )
{
    if ( !grid_delta || !grid_min || !grid_max || !num_bin )
        return fail( nullptr, CABANA_AMD_ERROR_INVALID_ARGUMENT, "NULL argument" );
    CartGridDev g;
    int64_t ncell;
    if ( !cart_init( grid_delta, grid_min, grid_max, g, ncell ) )
        return fail( nullptr, CABANA_AMD_ERROR_INVALID_ARGUMENT,
                     "invalid Cartesian grid" );
    for ( int d = 0; d < 3; ++d )
        num_bin[d] = g.nx[d];
    return CABANA_AMD_OK;
}

extern "C" cabana_amd_status cabana_amd_linked_cell_build(
    cabana_amd_context* ctx, const cabana_amd_particle_field* positions,
    int64_t begin, int64_t end, const double grid_delta[3],
    const double grid_min[3], const double grid_max[3], int32_t* counts,
    int64_t* offsets, int64_t* permutes, cabana_amd_stream stream )
{
    TraceRange trace_range( "Cabana::LinkedCellList::build" );
    if ( !ctx )
        return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT, "ctx is NULL" );
    if ( !positions || !grid_delta || !grid_min || !grid_max || !counts ||
         !offsets || ( !permutes && end > begin ) )
        return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT, "NULL argument" );
    if ( begin < 0 || end < begin || end > positions->num_particle )
        return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT,
                     "begin/end outside the particle range" );
    PFieldDev pos;
    CA_CHECK( make_pfield( ctx, *positions, 3, CABANA_AMD_FLOAT64, end, false,
                           "positions", pos ) );
    CartGridDev g;
    int64_t ncell;
    if ( !cart_init( grid_delta, grid_min, grid_max, g, ncell ) )
        return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT,
                     "invalid Cartesian grid" );
    if ( ncell >= INT32_MAX )
        return fail( ctx, CABANA_AMD_ERROR_UNSUPPORTED,
                     "more than 2^31 - 2 bins" );
    hipStream_t s = static_cast<hipStream_t>( stream );
    const int64_t np = end - begin;
    // Part tables derived from the earlier contents of counts/offsets are
    // retired now, before anything is written.
    bump_generation( counts, offsets );
    // Keys are 0..ncell-1 and ncell for particles outside the grid.
    int key_bits = 0;
    while ( ( ncell >> key_bits ) != 0 )
        ++key_bits;
    int digit_bits;
    lcl_sort_passes( key_bits, &digit_bits );
    const int64_t nhist = ( (int64_t)1 << digit_bits ) * lcl_sort_tiles( np );
    const int64_t nblock = ( nhist + 1023 ) / 1024 + 1;
    const int64_t nkey = np + 64;
    CA_HIP( ctx, ctx->ensure( SLOT_CELL_OF, sizeof( int32_t ) * 4 * nkey ),
            "scratch" );
    CA_HIP( ctx,
            ctx->ensure( SLOT_CURSOR,
                         ( sizeof( int64_t ) + sizeof( int32_t ) ) * ( nhist + 64 ) ),
            "scratch" );
    CA_HIP( ctx, ctx->ensure( SLOT_SCAN, sizeof( int64_t ) * nblock ),
            "scratch" );
    int32_t* keys = static_cast<int32_t*>( ctx->scratch[SLOT_CELL_OF] );
    int64_t* digit_base = static_cast<int64_t*>( ctx->scratch[SLOT_CURSOR] );
    int32_t* hist = reinterpret_cast<int32_t*>( digit_base + nhist + 64 );
    int64_t* sums = static_cast<int64_t*>( ctx->scratch[SLOT_SCAN] );
    CA_HIP( ctx,
            launch_fill_i32( *ctx, reinterpret_cast<int32_t*>( ctx->bad ), 2, 0,
                             s ),
            "fill" );
    CA_HIP( ctx,
            launch_lcl_keys( *ctx, pos, begin, end, g, (int32_t)ncell, keys,
                             ctx->bad, s ),
            "bin keys" );
    int32_t* sorted = keys;
    CA_HIP( ctx,
            launch_lcl_sort( *ctx, np, key_bits, keys, keys + nkey,
                             keys + 2 * nkey, keys + 3 * nkey, hist, digit_base,
                             sums, nblock, permutes, begin, &sorted, s ),
            "bin sort" );
    CA_HIP( ctx, launch_lcl_bounds( *ctx, sorted, np, ncell, counts, offsets, s ),
            "bin bounds" );
    // The reference fences after each phase; one wait here reports
    // particles outside the grid.
    unsigned long long bad = 0;
    CA_HIP( ctx,
            hipMemcpyAsync( &bad, ctx->bad, sizeof( bad ),
                            hipMemcpyDeviceToHost, s ),
            "hipMemcpyAsync" );
    CA_HIP( ctx, hipStreamSynchronize( s ), "hipStreamSynchronize" );
    if ( bad )
        return fail( ctx, CABANA_AMD_ERROR_OUT_OF_DOMAIN,
                     std::to_string( bad ) +
                         " particles lie outside the binning grid" );
    return CABANA_AMD_OK;
}

extern "C" cabana_amd_status cabana_amd_permute(
    cabana_amd_context* ctx, const int64_t* permutes, int64_t begin,
    int64_t end, const cabana_amd_particle_field* field,
    cabana_amd_stream stream )
{
    TraceRange trace_range( "Cabana::permute" );
    if ( !ctx )
        return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT, "ctx is NULL" );
    if ( !field || ( !permutes && end > begin ) )
        return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT, "NULL argument" );
    if ( begin < 0 || end < begin || end > field->num_particle )
        return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT,
                     "begin/end outside the particle range" );
    const int es = dtype_size( field->dtype );
    if ( !es )
        return fail( ctx, CABANA_AMD_ERROR_UNSUPPORTED, "unknown dtype" );
    PFieldDev f;
    CA_CHECK( make_pfield( ctx, *field, -1, -1, end, false, "field", f ) );
    hipStream_t s = static_cast<hipStream_t>( stream );
    const int64_t np = end - begin;
    CA_HIP( ctx,
            ctx->ensure( SLOT_PERMUTE, (size_t)es * f.ncomp * ( np + 1 ) ),
            "scratch" );
    void* scratch = ctx->scratch[SLOT_PERMUTE];
    CA_HIP( ctx,
            launch_permute_gather( *ctx, f, es, permutes, begin, end, scratch, s ),
            "permute gather" );
    CA_HIP( ctx, launch_permute_copy_back( *ctx, f, es, begin, end, scratch, s ),
            "permute copy back" );
    return CABANA_AMD_OK;
}

extern "C" cabana_amd_status cabana_amd_permute_members(
    cabana_amd_context* ctx, const int64_t* permutes, int64_t begin, int64_t end,
    const cabana_amd_particle_field* fields, int32_t num_fields,
    cabana_amd_stream stream )
{
    TraceRange trace_range( "Cabana::permute" );
    if ( !ctx )
        return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT, "ctx is NULL" );
    if ( num_fields < 0 || ( num_fields > 0 && !fields ) || ( !permutes && end > begin ) )
        return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT, "NULL argument" );
    std::vector<PFieldDev> f( (size_t)num_fields );
    std::vector<int> es( (size_t)num_fields );
    size_t row_bytes = 0;
    for ( int k = 0; k < num_fields; ++k )
    {
        if ( begin < 0 || end < begin || end > fields[k].num_particle )
            return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT,
                         "begin/end outside the particle range" );
        es[k] = dtype_size( fields[k].dtype );
        if ( !es[k] )
            return fail( ctx, CABANA_AMD_ERROR_UNSUPPORTED, "unknown dtype" );
        CA_CHECK( make_pfield( ctx, fields[k], -1, -1, end, false, "field", f[k] ) );
        row_bytes += (size_t)es[k] * f[k].ncomp;
    }
    if ( end <= begin || num_fields == 0 )
        return CABANA_AMD_OK;
    hipStream_t s = static_cast<hipStream_t>( stream );
    CA_HIP( ctx, ctx->ensure( SLOT_PERMUTE, row_bytes * (size_t)( end - begin + 1 ) ),
            "scratch" );
    CA_HIP( ctx,
            launch_permute_members( *ctx, f.data(), es.data(), num_fields, permutes, begin,
                                    end, ctx->scratch[SLOT_PERMUTE], s ),
            "permute members" );
    return CABANA_AMD_OK;
}

// ---------------------------------------------------------------------------
// BOV output (one rank)
// ---------------------------------------------------------------------------

extern "C" cabana_amd_status cabana_amd_bov_write_time_step(
    cabana_amd_context* ctx, const cabana_amd_mesh* mesh, const cabana_amd_grid_field* grid,
    int32_t entity, const char* label, const char* prefix, int32_t time_step_index,
    double time, int32_t gather_array, cabana_amd_stream stream )
{
    TraceRange trace_range( "Cabana::Grid::BovWriter::writeTimeStep" );
    if ( !ctx )
        return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT, "ctx is NULL" );
    if ( !grid || !label || !prefix )
        return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT, "NULL argument" );
    if ( entity != CABANA_AMD_NODE && entity != CABANA_AMD_CELL )
        return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT, "entity must be NODE or CELL" );
    if ( !mesh )
        return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT, "mesh is NULL" );
    // The array's entity is the call's argument (the writer's centering).
    cabana_amd_mesh me = *mesh;
    me.entity = entity;
    MeshDev m;
    CA_CHECK( make_mesh( ctx, &me, m ) );
    bool fp64 = true;
    CA_CHECK( precision_of( ctx, grid->dtype, fp64 ) );
    const bool node = entity == CABANA_AMD_NODE;
    GFieldDev g;
    CA_CHECK( make_gfield( ctx, m, *grid, -1, grid->dtype, false, "grid", g ) );
    hipStream_t s = static_cast<hipStream_t>( stream );
    // writeTimeStep (Cabana_Grid_BovWriter.hpp:180-338) on one rank, the
    // last block of every dimension: global extents are the cell counts
    // (+1 for nodes); the written block starts at the first owned entity
    // and covers all of them, including for nodes the closing node of a
    // periodic dimension, which the width-0 node halo gather fills from its
    // periodic owner first when gather_array is set.
    int n[3];
    for ( int d = 0; d < 3; ++d )
        n[d] = m.num_cell[d] + ( node ? 1 : 0 );
    bool any_periodic = false;
    for ( int d = 0; d < 3; ++d )
        any_periodic = any_periodic || m.periodic[d];
    if ( node && gather_array && any_periodic )
        CA_HIP( ctx, launch_halo_gather( *ctx, fp64, m, g, 0, s ), "halo gather" );
    const int dofs = grid->extent[3];
    const size_t es = fp64 ? 8 : 4;
    const size_t count = (size_t)n[0] * n[1] * n[2] * dofs;
    CA_HIP( ctx, ctx->ensure( SLOT_PERMUTE, es * count + 16 ), "scratch" );
    CA_HIP( ctx, launch_bov_pack( *ctx, fp64, g, m.halo, n, ctx->scratch[SLOT_PERMUTE], s ),
            "bov pack" );
    std::vector<char> host( es * count );
    CA_HIP( ctx,
            hipMemcpyAsync( host.data(), ctx->scratch[SLOT_PERMUTE], es * count,
                            hipMemcpyDeviceToHost, s ),
            "hipMemcpyAsync" );
    CA_HIP( ctx, hipStreamSynchronize( s ), "hipStreamSynchronize" );

    std::stringstream file_name;
    file_name << prefix << "_" << std::setfill( '0' ) << std::setw( 6 ) << time_step_index;
    const std::string data_file_name = file_name.str() + ".dat";
    {
        std::ofstream data( data_file_name, std::ios::out | std::ios::binary | std::ios::trunc );
        if ( !data )
            return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT,
                         "cannot open " + data_file_name );
        data.write( host.data(), (std::streamsize)host.size() );
        if ( !data )
            return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT,
                         "cannot write " + data_file_name );
    }
    // The header, field by field and in the stream formatting of
    // Cabana_Grid_BovWriter.hpp:276-337.
    const std::string header_file_name = file_name.str() + ".bov";
    std::fstream header;
    header.open( header_file_name, std::fstream::out );
    if ( !header )
        return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT,
                     "cannot open " + header_file_name );
    header << "TIME: " << time << std::endl;
    header << "DATA_FILE: " << data_file_name << std::endl;
    header << "DATA_SIZE: ";
    for ( int d = 0; d < 3; ++d )
        header << (long)n[d] << " ";
    header << std::endl;
    header << "DATA_FORMAT: " << ( fp64 ? "DOUBLE" : "FLOAT" ) << std::endl;
    header << "VARIABLE: " << label << std::endl;
    header << "DATA_ENDIAN: LITTLE" << std::endl;
    header << "CENTERING: " << ( node ? "nodal" : "zonal" ) << std::endl;
    header << "BRICK_ORIGIN: ";
    for ( int d = 0; d < 3; ++d )
        header << mesh->low_corner[d] << " ";
    header << std::endl;
    header << "BRICK_SIZE: ";
    for ( int d = 0; d < 3; ++d )
        header << m.num_cell[d] * mesh->cell_size[d] << " ";
    header << std::endl;
    header << "DATA_COMPONENTS: " << (long)dofs << std::endl;
    header.close();
    if ( !header )
        return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT,
                     "cannot write " + header_file_name );
    return CABANA_AMD_OK;
}

// ---------------------------------------------------------------------------
// Particle migration (one rank)
// ---------------------------------------------------------------------------

namespace
{
// Reads one device counter back (the call waits for the stream, as the
// reference's migrateCount waits for its reduction).
cabana_amd_status read_counter( cabana_amd_context* ctx, hipStream_t s,
                                unsigned long long& v )
{
    CA_HIP( ctx, hipMemcpyAsync( &v, ctx->bad, sizeof( v ), hipMemcpyDeviceToHost, s ),
            "hipMemcpyAsync" );
    CA_HIP( ctx, hipStreamSynchronize( s ), "hipStreamSynchronize" );
    return CABANA_AMD_OK;
}
} // namespace

extern "C" cabana_amd_status cabana_amd_particle_migrate(
    cabana_amd_context* ctx, const cabana_amd_mesh* mesh,
    const cabana_amd_particle_field* positions, const cabana_amd_particle_field* members,
    int32_t num_members, int32_t min_halo_width, int32_t force_migrate,
    int64_t* num_particle, int32_t* migrated, cabana_amd_stream stream )
{
    TraceRange trace_range( "Cabana::Grid::particleMigrate" );
    if ( !ctx )
        return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT, "ctx is NULL" );
    if ( !positions || !num_particle || !migrated || num_members < 0 ||
         ( num_members > 0 && !members ) )
        return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT, "NULL argument" );
    if ( min_halo_width < 0 )
        return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT, "negative minimum halo width" );
    MeshDev m;
    CA_CHECK( make_mesh( ctx, mesh, m ) );
    const int64_t n = positions->num_particle;
    if ( n <= 0 )
    {
        // No particle is outside the band; a forced migration moves none.
        *num_particle = 0;
        *migrated = force_migrate ? 1 : 0;
        return CABANA_AMD_OK;
    }
    PFieldDev pos;
    CA_CHECK( make_pfield( ctx, *positions, 3, CABANA_AMD_FLOAT64, n, false, "positions",
                           pos ) );
    std::vector<PFieldDev> f( (size_t)num_members );
    std::vector<int> es( (size_t)num_members );
    for ( int k = 0; k < num_members; ++k )
    {
        if ( members[k].num_particle < n )
            return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT,
                         "a member holds fewer particles than the positions" );
        es[k] = dtype_size( members[k].dtype );
        if ( !es[k] )
            return fail( ctx, CABANA_AMD_ERROR_UNSUPPORTED, "unknown dtype" );
        CA_CHECK( make_pfield( ctx, members[k], -1, -1, n, false, "member", f[k] ) );
    }
    // LocalMesh (Cabana_Grid_LocalMesh.hpp:61-107) and GlobalMesh corners of
    // one rank, in the same FP64 operations.
    MigrateBoundsDev b;
    for ( int d = 0; d < 3; ++d )
    {
        const double cell = mesh->cell_size[d];
        const double own_lo = mesh->low_corner[d] + cell * 0;
        const double own_hi = mesh->low_corner[d] + cell * ( 0 + m.num_cell[d] );
        const double ghost_lo = own_lo - mesh->halo_cell_width * cell;
        const double ghost_hi = own_hi + mesh->halo_cell_width * cell;
        b.band_lo[d] = ghost_lo + min_halo_width * cell;
        b.band_hi[d] = ghost_hi - min_halo_width * cell;
        b.own_lo[d] = own_lo;
        b.own_hi[d] = own_hi;
        b.global_lo[d] = mesh->low_corner[d];
        b.global_hi[d] = mesh->high_corner[d];
        b.extent[d] = mesh->high_corner[d] - mesh->low_corner[d];
        b.periodic[d] = m.periodic[d];
    }
    hipStream_t s = static_cast<hipStream_t>( stream );
    *num_particle = n;
    *migrated = 0;
    if ( !force_migrate )
    {
        CA_HIP( ctx, launch_fill_i32( *ctx, reinterpret_cast<int32_t*>( ctx->bad ), 2, 0, s ),
                "fill" );
        CA_HIP( ctx, launch_migrate_count( *ctx, pos, n, b, ctx->bad, s ), "migrate count" );
        unsigned long long outside = 0;
        CA_CHECK( read_counter( ctx, s, outside ) );
        if ( outside == 0 )
            return CABANA_AMD_OK;
    }
    *migrated = 1;
    const int64_t nblock = ( n + 1023 ) / 1024 + 1;
    CA_HIP( ctx, ctx->ensure( SLOT_CELL_OF, sizeof( int32_t ) * ( n + 64 ) ), "scratch" );
    CA_HIP( ctx, ctx->ensure( SLOT_CURSOR, sizeof( int64_t ) * 2 * ( n + 64 ) ), "scratch" );
    CA_HIP( ctx, ctx->ensure( SLOT_SCAN, sizeof( int64_t ) * nblock ), "scratch" );
    int32_t* keep = static_cast<int32_t*>( ctx->scratch[SLOT_CELL_OF] );
    int64_t* rank = static_cast<int64_t*>( ctx->scratch[SLOT_CURSOR] );
    int64_t* idx = rank + n + 64;
    int64_t* sums = static_cast<int64_t*>( ctx->scratch[SLOT_SCAN] );
    CA_HIP( ctx, launch_fill_i32( *ctx, reinterpret_cast<int32_t*>( ctx->bad ), 2, 0, s ),
            "fill" );
    CA_HIP( ctx, launch_migrate_shift( *ctx, pos, n, b, keep, ctx->bad, s ), "migrate shift" );
    unsigned long long dropped = 0;
    CA_CHECK( read_counter( ctx, s, dropped ) );
    if ( dropped == 0 || num_members == 0 )
    {
        *num_particle = n - (int64_t)dropped;
        return CABANA_AMD_OK;
    }
    // Distributor::migrate keeps the staying particles in their order: a
    // stable compaction of every member.
    const int64_t kept = n - (int64_t)dropped;
    CA_HIP( ctx, launch_exclusive_scan( *ctx, keep, rank, n, sums, nblock, s ), "scan" );
    CA_HIP( ctx, launch_compact_index( *ctx, keep, rank, n, idx, s ), "compact index" );
    if ( kept > 0 )
    {
        size_t row_bytes = 0;
        for ( int k = 0; k < num_members; ++k )
            row_bytes += (size_t)es[k] * f[k].ncomp;
        CA_HIP( ctx, ctx->ensure( SLOT_PERMUTE, row_bytes * (size_t)( kept + 1 ) ), "scratch" );
        CA_HIP( ctx,
                launch_permute_members( *ctx, f.data(), es.data(), num_members, idx, 0, kept,
                                        ctx->scratch[SLOT_PERMUTE], s ),
                "permute members" );
    }
    *num_particle = kept;
    return CABANA_AMD_OK;
}

// ---------------------------------------------------------------------------
// Profiling
// ---------------------------------------------------------------------------

extern "C" cabana_amd_status cabana_amd_probe_lds_add( cabana_amd_context* ctx, int32_t fp64,
                                                      int32_t iters,
                                                      double* cycles_per_wave_op )
{
    if ( !ctx || !cycles_per_wave_op || iters < 1 || iters > ( 1 << 24 ) )
        return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT,
                     "need ctx, an output and 1 <= iters <= 2^24" );
    int dev = 0, cus = 0, khz = 0;
    CA_HIP( ctx, hipGetDevice( &dev ), "get device" );
    CA_HIP( ctx, hipDeviceGetAttribute( &cus, hipDeviceAttributeMultiprocessorCount, dev ),
            "CU count" );
    CA_HIP( ctx, hipDeviceGetAttribute( &khz, hipDeviceAttributeClockRate, dev ), "clock rate" );
    const int groups = 4 * cus; // 16 waves per CU
    void* out = nullptr;
    CA_HIP( ctx, hipMalloc( &out, sizeof( double ) * groups ), "probe buffer" );
    float ms = 0.0f;
    const hipError_t e = probe_lds_add( fp64 != 0, iters, groups, out, nullptr, ms );
    (void)hipFree( out );
    CA_HIP( ctx, e, "LDS probe" );
    // Each CU runs 16 waves of iters wave-wide adds.
    *cycles_per_wave_op = (double)ms * 1e-3 * (double)khz * 1e3 / ( 16.0 * iters );
    return CABANA_AMD_OK;
}

extern "C" cabana_amd_status cabana_amd_set_deterministic( cabana_amd_context* ctx,
                                                           int32_t enable )
{
    if ( !ctx )
        return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT, "ctx is NULL" );
    ctx->deterministic = enable != 0;
    return CABANA_AMD_OK;
}

extern "C" cabana_amd_status cabana_amd_set_auto_binning( cabana_amd_context* ctx,
                                                          int32_t enable )
{
    if ( !ctx )
        return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT, "ctx is NULL" );
    ctx->auto_binning = enable != 0;
    return CABANA_AMD_OK;
}

extern "C" cabana_amd_status cabana_amd_profile_enable( cabana_amd_context* ctx,
                                                        int32_t enable )
{
    if ( !ctx )
        return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT, "ctx is NULL" );
    ctx->profiling = enable != 0;
    return CABANA_AMD_OK;
}

extern "C" cabana_amd_status cabana_amd_profile_reset( cabana_amd_context* ctx )
{
    if ( !ctx )
        return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT, "ctx is NULL" );
    CA_HIP( ctx, ctx->drain(), "event synchronize" );
    ctx->stats.clear();
    ctx->order.clear();
    return CABANA_AMD_OK;
}

extern "C" cabana_amd_status cabana_amd_profile_count( cabana_amd_context* ctx,
                                                       int32_t* count )
{
    if ( !ctx || !count )
        return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT, "NULL argument" );
    CA_HIP( ctx, ctx->drain(), "event synchronize" );
    *count = (int32_t)ctx->order.size();
    return CABANA_AMD_OK;
}

extern "C" cabana_amd_status cabana_amd_profile_entry(
    cabana_amd_context* ctx, int32_t index, char* name, int32_t name_capacity,
    int64_t* launches, double* total_ms )
{
    if ( !ctx )
        return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT, "ctx is NULL" );
    if ( index < 0 || index >= (int32_t)ctx->order.size() )
        return fail( ctx, CABANA_AMD_ERROR_INVALID_ARGUMENT,
                     "profile index out of range" );
    const std::string& nm = ctx->order[index];
    const Stat& st = ctx->stats[nm];
    if ( name && name_capacity > 0 )
    {
        std::snprintf( name, (size_t)name_capacity, "%s", nm.c_str() );
    }
    if ( launches )
        *launches = st.launches;
    if ( total_ms )
        *total_ms = st.ms;
    return CABANA_AMD_OK;
}
