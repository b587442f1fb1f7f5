// cabana_amd_p2g_no_alloc: p2g for callers that cannot allocate device
// memory (code under HIP graph capture, serving loops, harnesses that forbid
// hipMalloc).  Its only scratch, the brick part table of a tiled scatter,
// lives in a fixed device array of this file; the library allocates nothing
// during the call.
#include "ca_internal.hpp"

#include <map>
#include <mutex>

// One copy per device, held until the process ends.  Host code never uses
// this name as an address: each device's copy is looked up once, with that
// device current, and that pointer is what the kernels receive.
static __device__ __align__( 256 ) unsigned char scratch[CABANA_AMD_NO_ALLOC_SCRATCH_BYTES];

namespace
{
std::mutex g_mutex;
std::map<int, void*> g_address; // device -> its copy of `scratch`

// Address of the current device's copy (looked up on the first call on the
// device); nullptr when the lookup fails.
void* scratch_of_current_device( int& device )
{
    if ( hipGetDevice( &device ) != hipSuccess )
        return nullptr;
    std::lock_guard<std::mutex> lock( g_mutex );
    const auto it = g_address.find( device );
    if ( it != g_address.end() )
        return it->second;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress( (void**)&ptr, HIP_SYMBOL( scratch ) );
    if ( e != hipSuccess || !ptr )
    {
        (void)hipGetLastError();
        return nullptr;
    }
    g_address[device] = ptr;
    return ptr;
}
} // namespace

extern "C" cabana_amd_status cabana_amd_p2g_no_alloc(
    cabana_amd_context* ctx, const cabana_amd_mesh* mesh, int32_t order,
    const cabana_amd_particle_field* points, int32_t num_terms,
    const cabana_amd_p2g_term* terms, int32_t halo_width,
    const cabana_amd_binning* bins, cabana_amd_stream stream )
{
    int device = -1;
    void* ptr = scratch_of_current_device( device );
    if ( !ptr )
        return cabana_amd::fail_with( ctx, CABANA_AMD_ERROR_DEVICE,
                                      "hipGetSymbolAddress of the no-allocation scratch failed" );
    return cabana_amd::p2g_fixed_scratch( ctx, mesh, order, points, num_terms, terms, halo_width,
                                          bins, stream, ptr, CABANA_AMD_NO_ALLOC_SCRATCH_BYTES );
}
