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
#include <type_traits>
#include <utility>

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
// wave intrinsics, for one lane at a time: its own value and its own vote (code that takes a
// wave-uniform branch then takes it with the lane's own value, which is what it computes)
#define __builtin_amdgcn_readfirstlane( x ) ( x )
inline bool __all( int predicate )
{
	return predicate != 0;
}
// a thread that polls (s_sleep in a wait loop) gives the other threads of a cooperative launch
// their turn
void emu_yield();
#define __builtin_amdgcn_s_sleep( n ) emu_yield()
// one host thread runs every kernel thread in turn, so plain loads and stores are the device's
// atomics and fences
#define __threadfence() ( (void)0 )
template <typename T>
inline T __hip_atomic_load( const T* p, int, int )
{
	return *p;
}
template <typename T>
inline void __hip_atomic_store( T* p, T v, int, int )
{
	*p = v;
}
template <typename T>
inline T __hip_atomic_fetch_add( T* p, T v, int, int )
{
	T old = *p;
	*p = old + v;
	return old;
}
#define __HIP_MEMORY_SCOPE_AGENT 3

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
inline int atomicSub( int* p, int v )
{
	int old = *p;
	*p = old - v;
	return old;
}
inline int atomicExch( int* p, int v )
{
	int old = *p;
	*p = v;
	return old;
}
inline int atomicCAS( int* p, int compare, int v )
{
	int old = *p;
	if ( old == compare )
		*p = v;
	return old;
}
inline unsigned long long atomicCAS( unsigned long long* p, unsigned long long compare, unsigned long long v )
{
	unsigned long long old = *p;
	if ( old == compare )
		*p = v;
	return old;
}
inline uint32_t atomicOr( uint32_t* p, uint32_t v )
{
	uint32_t old = *p;
	*p = old | v;
	return old;
}
inline uint32_t atomicAnd( uint32_t* p, uint32_t v )
{
	uint32_t old = *p;
	*p = old & v;
	return old;
}

// ---- runtime API ----
typedef enum hipError_t
{
	hipSuccess = 0,
	hipErrorInvalidValue = 1,
	hipErrorOutOfMemory = 2,
	hipErrorNoDevice = 100,
	hipErrorNotSupported = 801,
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

// stream capture is not emulated: a capture is refused, so callers launch directly
struct ihipGraph_t;
struct ihipGraphExec_t;
typedef ihipGraph_t* hipGraph_t;
typedef ihipGraphExec_t* hipGraphExec_t;
typedef enum hipStreamCaptureMode
{
	hipStreamCaptureModeGlobal = 0,
	hipStreamCaptureModeThreadLocal = 1,
	hipStreamCaptureModeRelaxed = 2,
} hipStreamCaptureMode;
hipError_t hipStreamBeginCapture( hipStream_t s, hipStreamCaptureMode mode );
hipError_t hipStreamEndCapture( hipStream_t s, hipGraph_t* graph );
hipError_t hipGraphInstantiate( hipGraphExec_t* exec, hipGraph_t graph, void* errorNode, char* log, size_t logSize );
hipError_t hipGraphLaunch( hipGraphExec_t exec, hipStream_t s );
hipError_t hipGraphExecDestroy( hipGraphExec_t exec );
hipError_t hipGraphDestroy( hipGraph_t graph );
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
hipError_t hipStreamDestroy( hipStream_t s );
hipError_t hipStreamWaitEvent( hipStream_t s, hipEvent_t e, unsigned flags );
hipError_t hipEventCreate( hipEvent_t* e );
hipError_t hipEventCreateWithFlags( hipEvent_t* e, unsigned flags );
hipError_t hipEventRecord( hipEvent_t e, hipStream_t s );
hipError_t hipEventElapsedTime( float* ms, hipEvent_t a, hipEvent_t b );
hipError_t hipEventDestroy( hipEvent_t e );
hipError_t hipFuncSetAttribute( const void* f, hipFuncAttribute attr, int value );
typedef enum hipDeviceAttribute_t
{
	hipDeviceAttributeCooperativeLaunch = 1,
	hipDeviceAttributeMultiprocessorCount = 2,
} hipDeviceAttribute_t;
hipError_t hipDeviceGetAttribute( int* value, hipDeviceAttribute_t attr, int device );
hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor( int* blocks, const void* f, int blockSize, size_t sharedBytes );
hipError_t hipGetLastError();
const char* hipGetErrorString( hipError_t e );

// dynamic shared memory of the running workgroup (checked against the launch's size)
void emu_launch( dim3 grid, dim3 block, size_t sharedBytes, const std::function<void()>& body );

#define hipLaunchKernelGGL( kernel, grid, block, shared, stream, ... ) \
	emu_launch( dim3( grid ), dim3( block ), ( shared ), [&]() { kernel( __VA_ARGS__ ); } )

// A cooperative launch: every workgroup's threads are live at once (fibers of one scheduler, each
// workgroup with its own barrier), so a grid-wide barrier built from atomics and polling works.
void emu_launch_cooperative( dim3 grid, dim3 block, size_t sharedBytes, const std::function<void()>& body );
template <typename... Args, size_t... I>
inline void emu_call_kernel( void ( *k )( Args... ), void** args, std::index_sequence<I...> )
{
	k( *static_cast<typename std::remove_reference<Args>::type*>( args[I] )... );
}
template <typename... Args>
inline hipError_t hipLaunchCooperativeKernel( void ( *k )( Args... ), dim3 grid, dim3 block, void** args, unsigned sharedBytes,
											  hipStream_t )
{
	emu_launch_cooperative( grid, block, sharedBytes, [&]() { emu_call_kernel( k, args, std::index_sequence_for<Args...>{} ); } );
	return hipSuccess;
}
