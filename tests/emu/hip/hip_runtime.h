// CPU stand-in for the HIP runtime and kernel language, used only to build the library's host
// driver AND its kernels as ordinary host code for tests on machines without a GPU (see
// tests/emu/emu_runtime.cpp). Device memory is host heap memory, so an out-of-bounds kernel
// access becomes a heap overflow that AddressSanitizer reports with the kernel's stack. Streams
// are synchronous; a kernel launch runs every workgroup in turn, each workgroup's threads as
// cooperative fibers that switch at __syncthreads().
#pragma once

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <functional>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__( ... )
#define __align__( n ) alignas( n )
// workgroups run one at a time on the launching thread, so its thread-local storage is the
// workgroup's shared memory
#define __shared__ thread_local

struct dim3
{
	uint32_t x, y, z;
	dim3( uint32_t x_ = 1, uint32_t y_ = 1, uint32_t z_ = 1 ) : x( x_ ), y( y_ ), z( z_ ) {}
};

struct uint3
{
	uint32_t x, y, z;
};

struct float2
{
	float x, y;
};
struct float4
{
	float x, y, z, w;
};
struct int2
{
	int x, y;
};
struct int4
{
	int x, y, z, w;
};
inline float2 make_float2( float x, float y )
{
	return float2{ x, y };
}
inline float4 make_float4( float x, float y, float z, float w )
{
	return float4{ x, y, z, w };
}
inline int2 make_int2( int x, int y )
{
	return int2{ x, y };
}
inline int4 make_int4( int x, int y, int z, int w )
{
	return int4{ x, y, z, w };
}

// The dynamic shared memory of the library's kernels (their one `extern __shared__` array,
// b2mi::lds), defined here as an inline variable so every translation unit accesses it
// directly: an `extern thread_local` array defined in another unit is reached through a TLS
// wrapper function that the defining unit need not emit.
namespace b2mi
{
constexpr size_t kEmuSharedBytes = 160 * 1024;
inline thread_local float4 lds[kEmuSharedBytes / sizeof( float4 )];
} // namespace b2mi

// ---- execution state of the running kernel thread ----
struct EmuThreadState
{
	uint3 tid;
};
extern thread_local EmuThreadState* emu_current;
extern thread_local uint3 emu_blockIdx;
extern thread_local dim3 emu_blockDim, emu_gridDim;
#define threadIdx ( emu_current->tid )
#define blockIdx ( emu_blockIdx )
#define blockDim ( emu_blockDim )
#define gridDim ( emu_gridDim )

void emu_syncthreads();
#define __syncthreads() emu_syncthreads()

inline int atomicAdd( int* p, int v )
{
	int old = *p;
	*p = old + v;
	return old;
}
inline unsigned atomicAdd( unsigned* p, unsigned v )
{
	unsigned old = *p;
	*p = old + v;
	return old;
}
inline float atomicAdd( float* p, float v )
{
	float old = *p;
	*p = old + v;
	return old;
}

// ---- runtime API ----
typedef enum hipError_t
{
	hipSuccess = 0,
	hipErrorInvalidValue = 1,
	hipErrorOutOfMemory = 2,
	hipErrorNoDevice = 100,
} hipError_t;

typedef enum hipMemcpyKind
{
	hipMemcpyHostToHost = 0,
	hipMemcpyHostToDevice = 1,
	hipMemcpyDeviceToHost = 2,
	hipMemcpyDeviceToDevice = 3,
	hipMemcpyDefault = 4,
} hipMemcpyKind;

typedef enum hipFuncAttribute
{
	hipFuncAttributeMaxDynamicSharedMemorySize = 8,
} hipFuncAttribute;

struct ihipStream_t;
struct ihipEvent_t;
typedef ihipStream_t* hipStream_t;
typedef ihipEvent_t* hipEvent_t;

#define hipStreamNonBlocking 1u
#define hipEventDisableTiming 2u
#define hipHostMallocDefault 0u

hipError_t hipGetDeviceCount( int* count );
hipError_t hipGetDevice( int* device );
// a __device__ variable is an ordinary host variable here, so its device address is its address
hipError_t hipGetSymbolAddress( void** p, const void* symbol );
#define HIP_SYMBOL( x ) ( &( x ) )
hipError_t hipMalloc( void** p, size_t bytes );
hipError_t hipFree( void* p );
hipError_t hipHostMalloc( void** p, size_t bytes, unsigned flags );
hipError_t hipHostFree( void* p );
hipError_t hipMemcpyAsync( void* dst, const void* src, size_t bytes, hipMemcpyKind kind, hipStream_t s );
hipError_t hipMemsetAsync( void* dst, int value, size_t bytes, hipStream_t s );
hipError_t hipStreamSynchronize( hipStream_t s );
hipError_t hipStreamCreateWithFlags( hipStream_t* s, unsigned flags );
hipError_t hipStreamWaitEvent( hipStream_t s, hipEvent_t e, unsigned flags );
hipError_t hipEventCreate( hipEvent_t* e );
hipError_t hipEventCreateWithFlags( hipEvent_t* e, unsigned flags );
hipError_t hipEventRecord( hipEvent_t e, hipStream_t s );
hipError_t hipEventElapsedTime( float* ms, hipEvent_t a, hipEvent_t b );
hipError_t hipEventDestroy( hipEvent_t e );
hipError_t hipFuncSetAttribute( const void* f, hipFuncAttribute attr, int value );
hipError_t hipGetLastError();
const char* hipGetErrorString( hipError_t e );

// dynamic shared memory of the running workgroup (checked against the launch's size)
void emu_launch( dim3 grid, dim3 block, size_t sharedBytes, const std::function<void()>& body );

#define hipLaunchKernelGGL( kernel, grid, block, shared, stream, ... ) \
	emu_launch( dim3( grid ), dim3( block ), ( shared ), [&]() { kernel( __VA_ARGS__ ); } )
