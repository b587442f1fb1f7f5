// LDS atomic-add throughput probe: the measurement behind the scatter's
// FP64 LDS window (its adds are ds_add_f64; an FP32 window would use
// ds_add_f32).  Every lane of 16 waves per CU adds into its own LDS word,
// stepping through a 4 KiB (FP32) or 8 KiB (FP64) buffer so that a wave's
// 64 addresses are distinct and spread over the banks.
#include "ca_internal.hpp"

namespace cabana_amd
{
namespace
{

constexpr int PROBE_BLOCK = 256;
constexpr int PROBE_WORDS = 1024;

template <class T>
__global__ __launch_bounds__( PROBE_BLOCK ) void lds_add_probe_kernel( int iters, T* out )
{
    __shared__ T buf[PROBE_WORDS];
    for ( int i = threadIdx.x; i < PROBE_WORDS; i += PROBE_BLOCK )
        buf[i] = T( 0 );
    __syncthreads();
    const int base = threadIdx.x;
    for ( int it = 0; it < iters; ++it )
        atomicAdd( &buf[( base + 64 * it ) & ( PROBE_WORDS - 1 )], T( 1 ) );
    __syncthreads();
    if ( threadIdx.x == 0 )
        out[blockIdx.x] = buf[0];
}

} // namespace

// Times iters LDS atomic adds per lane over num_groups workgroups of 256
// threads (4 per CU when num_groups = 4 x CUs); returns the kernel time.
hipError_t probe_lds_add( bool fp64, int iters, int num_groups, void* out, hipStream_t s,
                          float& ms )
{
    hipEvent_t e0, e1;
    hipError_t e = hipEventCreate( &e0 );
    if ( e != hipSuccess )
        return e;
    e = hipEventCreate( &e1 );
    if ( e != hipSuccess )
    {
        (void)hipEventDestroy( e0 );
        return e;
    }
    (void)hipEventRecord( e0, s );
    if ( fp64 )
        hipLaunchKernelGGL( lds_add_probe_kernel<double>, dim3( num_groups ), dim3( PROBE_BLOCK ),
                            0, s, iters, static_cast<double*>( out ) );
    else
        hipLaunchKernelGGL( lds_add_probe_kernel<float>, dim3( num_groups ), dim3( PROBE_BLOCK ),
                            0, s, iters, static_cast<float*>( out ) );
    e = hipGetLastError();
    (void)hipEventRecord( e1, s );
    if ( e == hipSuccess )
        e = hipEventSynchronize( e1 );
    if ( e == hipSuccess )
        e = hipEventElapsedTime( &ms, e0, e1 );
    (void)hipEventDestroy( e0 );
    (void)hipEventDestroy( e1 );
    return e;
}

} // namespace cabana_amd
