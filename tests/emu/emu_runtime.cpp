// CPU runtime behind tests/emu/hip/hip_runtime.h: host-heap "device" memory, synchronous
// streams, and kernel launches that run each workgroup's threads as cooperative fibers
// (ucontext), switching at every __syncthreads(). Test infrastructure only: it lets the whole
// step pipeline, kernels included, run under AddressSanitizer on a machine without a GPU.
#include "hip/hip_runtime.h"

#include <chrono>
#include <cstdio>
#include <mutex>
#include <cstdlib>
#include <sys/mman.h>
#include <ucontext.h>
#include <vector>

#if defined( __SANITIZE_ADDRESS__ )
	#include <sanitizer/asan_interface.h>
	#include <sanitizer/common_interface_defs.h>
	#define EMU_ASAN 1
#else
	#define EMU_ASAN 0
#endif

thread_local EmuThreadState* emu_current = nullptr;
thread_local uint3 emu_blockIdx{ 0, 0, 0 };
thread_local dim3 emu_blockDim, emu_gridDim;

namespace
{

constexpr size_t kStackBytes = 512 * 1024;

struct Fiber
{
	EmuThreadState state;
	ucontext_t ctx;
	void* stack = nullptr;
	bool done = false;
	bool waiting = false; // at a workgroup barrier
	uint3 block{ 0, 0, 0 };
};

struct Launch
{
	const std::function<void()>* body = nullptr;
	bool cooperative = false;
	std::vector<Fiber> fibers;
	ucontext_t scheduler;
	Fiber* running = nullptr;
#if EMU_ASAN
	void* schedulerFakeStack = nullptr;
	const void* schedulerStackBottom = nullptr;
	size_t schedulerStackSize = 0;
#endif
};

thread_local Launch* g_launch = nullptr;
// fiber stacks for reuse, shared by launches from every host thread
std::vector<void*> g_stackPool;
std::mutex g_stackMutex;

void fiberEntry()
{
	Launch* L = g_launch;
#if EMU_ASAN
	__sanitizer_finish_switch_fiber( nullptr, &L->schedulerStackBottom, &L->schedulerStackSize );
#endif
	( *L->body )();
	L->running->done = true;
#if EMU_ASAN
	__sanitizer_start_switch_fiber( nullptr, L->schedulerStackBottom, L->schedulerStackSize );
#endif
	setcontext( &L->scheduler );
}

void* takeStack()
{
	std::lock_guard<std::mutex> lock( g_stackMutex );
	if ( !g_stackPool.empty() )
	{
		void* p = g_stackPool.back();
		g_stackPool.pop_back();
		return p;
	}
	void* p = mmap( nullptr, kStackBytes, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE, -1, 0 );
	if ( p == MAP_FAILED )
	{
		fprintf( stderr, "emu: cannot map a fiber stack\n" );
		abort();
	}
	return p;
}

void runFiber( Launch* L, Fiber* f )
{
	L->running = f;
	emu_current = &f->state;
#if EMU_ASAN
	__sanitizer_start_switch_fiber( &L->schedulerFakeStack, f->stack, kStackBytes );
#endif
	swapcontext( &L->scheduler, &f->ctx );
#if EMU_ASAN
	__sanitizer_finish_switch_fiber( L->schedulerFakeStack, nullptr, nullptr );
#endif
	emu_current = nullptr;
}

} // namespace

void emu_syncthreads()
{
	Launch* L = g_launch;
	Fiber* f = L->running;
	f->waiting = true;
#if EMU_ASAN
	void* fake = nullptr;
	__sanitizer_start_switch_fiber( &fake, L->schedulerStackBottom, L->schedulerStackSize );
#endif
	swapcontext( &f->ctx, &L->scheduler );
#if EMU_ASAN
	__sanitizer_finish_switch_fiber( fake, &L->schedulerStackBottom, &L->schedulerStackSize );
#endif
}

void emu_yield()
{
	Launch* L = g_launch;
	if ( L == nullptr || !L->cooperative )
		return; // one workgroup at a time: nothing else could run meanwhile
	Fiber* f = L->running;
	f->waiting = false;
#if EMU_ASAN
	void* fake = nullptr;
	__sanitizer_start_switch_fiber( &fake, L->schedulerStackBottom, L->schedulerStackSize );
#endif
	swapcontext( &f->ctx, &L->scheduler );
#if EMU_ASAN
	__sanitizer_finish_switch_fiber( fake, &L->schedulerStackBottom, &L->schedulerStackSize );
#endif
}

// Every thread of every workgroup is a fiber of one scheduler. A workgroup's barrier opens when
// all of its live threads wait at it; a polling thread (emu_yield) hands its turn on.
void emu_launch_cooperative( dim3 grid, dim3 block, size_t sharedBytes, const std::function<void()>& body )
{
	if ( sharedBytes != 0 || grid.y != 1 || grid.z != 1 || block.y != 1 || block.z != 1 )
	{
		fprintf( stderr, "emu: cooperative launches take one-dimensional grids and no dynamic shared memory\n" );
		abort();
	}
	const size_t threads = block.x, blocks = grid.x;
	if ( threads == 0 || threads > 1024 || blocks == 0 || blocks > 64 )
	{
		fprintf( stderr, "emu: cooperative launch of %zu x %zu threads\n", blocks, threads );
		abort();
	}
	Launch L;
	L.body = &body;
	L.cooperative = true;
	L.fibers.resize( threads * blocks );
	for ( size_t b = 0; b < blocks; ++b )
		for ( size_t t = 0; t < threads; ++t )
		{
			Fiber& f = L.fibers[b * threads + t];
			f.stack = takeStack();
			f.state.tid = uint3{ (uint32_t)t, 0, 0 };
			f.block = uint3{ (uint32_t)b, 0, 0 };
			f.done = false;
			f.waiting = false;
			getcontext( &f.ctx );
			f.ctx.uc_stack.ss_sp = f.stack;
			f.ctx.uc_stack.ss_size = kStackBytes;
			f.ctx.uc_link = nullptr;
			makecontext( &f.ctx, fiberEntry, 0 );
		}
	Launch* outer = g_launch;
	g_launch = &L;
	emu_gridDim = grid;
	emu_blockDim = block;
#if EMU_ASAN
	ASAN_POISON_MEMORY_REGION( reinterpret_cast<unsigned char*>( b2mi::lds ), b2mi::kEmuSharedBytes );
#endif
	for ( ;; )
	{
		// open the barriers whose workgroups are all there
		for ( size_t b = 0; b < blocks; ++b )
		{
			size_t live = 0, waiting = 0;
			for ( size_t t = 0; t < threads; ++t )
			{
				const Fiber& f = L.fibers[b * threads + t];
				live += f.done ? 0 : 1;
				waiting += ( !f.done && f.waiting ) ? 1 : 0;
			}
			if ( live > 0 && waiting == live )
			{
				if ( live != threads )
				{
					fprintf( stderr, "emu: %zu threads of workgroup %zu wait at a barrier that %zu others left\n", live, b,
							 threads - live );
					abort();
				}
				for ( size_t t = 0; t < threads; ++t )
					L.fibers[b * threads + t].waiting = false;
			}
		}
		size_t ran = 0, live = 0;
		for ( Fiber& f : L.fibers )
		{
			if ( f.done )
				continue;
			live += 1;
			if ( f.waiting )
				continue;
			emu_blockIdx = f.block;
			runFiber( &L, &f );
			ran += 1;
		}
		if ( live == 0 )
			break;
		if ( ran == 0 )
		{
			fprintf( stderr, "emu: cooperative launch deadlocked\n" );
			abort();
		}
	}
#if EMU_ASAN
	ASAN_UNPOISON_MEMORY_REGION( reinterpret_cast<unsigned char*>( b2mi::lds ), b2mi::kEmuSharedBytes );
#endif
	{
		std::lock_guard<std::mutex> lock( g_stackMutex );
		for ( Fiber& f : L.fibers )
			g_stackPool.push_back( f.stack );
	}
	g_launch = outer;
}

void emu_launch( dim3 grid, dim3 block, size_t sharedBytes, const std::function<void()>& body )
{
	if ( sharedBytes > b2mi::kEmuSharedBytes )
	{
		fprintf( stderr, "emu: launch asks for %zu bytes of shared memory\n", sharedBytes );
		abort();
	}
	size_t threads = (size_t)block.x * block.y * block.z;
	if ( threads == 0 || threads > 1024 )
	{
		fprintf( stderr, "emu: launch of %zu threads per workgroup\n", threads );
		abort();
	}
	Launch L;
	L.body = &body;
	L.fibers.resize( threads );
	for ( Fiber& f : L.fibers )
		f.stack = takeStack();
	Launch* outer = g_launch;
	g_launch = &L;
	emu_gridDim = grid;
	emu_blockDim = block;
	unsigned char* shared = reinterpret_cast<unsigned char*>( b2mi::lds );
	for ( uint32_t bz = 0; bz < grid.z; ++bz )
		for ( uint32_t by = 0; by < grid.y; ++by )
			for ( uint32_t bx = 0; bx < grid.x; ++bx )
			{
				emu_blockIdx = uint3{ bx, by, bz };
				// shared memory starts as NaN bytes; only the launch's size is addressable
#if EMU_ASAN
				ASAN_UNPOISON_MEMORY_REGION( shared, b2mi::kEmuSharedBytes );
#endif
				memset( shared, 0xff, b2mi::kEmuSharedBytes );
#if EMU_ASAN
				ASAN_POISON_MEMORY_REGION( shared + sharedBytes, b2mi::kEmuSharedBytes - sharedBytes );
#endif
				size_t k = 0;
				for ( uint32_t tz = 0; tz < block.z; ++tz )
					for ( uint32_t ty = 0; ty < block.y; ++ty )
						for ( uint32_t tx = 0; tx < block.x; ++tx, ++k )
						{
							Fiber& f = L.fibers[k];
							f.state.tid = uint3{ tx, ty, tz };
							f.done = false;
							f.waiting = false;
							getcontext( &f.ctx );
							f.ctx.uc_stack.ss_sp = f.stack;
							f.ctx.uc_stack.ss_size = kStackBytes;
							f.ctx.uc_link = nullptr;
							makecontext( &f.ctx, fiberEntry, 0 );
						}
				// rounds: every live thread runs to its next barrier (or its end)
				for ( ;; )
				{
					size_t live = 0, atBarrier = 0;
					for ( Fiber& f : L.fibers )
					{
						if ( f.done )
							continue;
						f.waiting = false;
						runFiber( &L, &f );
						if ( !f.done )
						{
							live += 1;
							atBarrier += f.waiting ? 1 : 0;
						}
					}
					if ( live == 0 )
						break;
					size_t doneCount = 0;
					for ( Fiber& f : L.fibers )
						doneCount += f.done ? 1 : 0;
					if ( doneCount != 0 )
					{
						fprintf( stderr, "emu: %zu threads of a workgroup wait at a barrier that %zu others left\n", live,
								 doneCount );
						abort();
					}
				}
			}
#if EMU_ASAN
	ASAN_UNPOISON_MEMORY_REGION( shared, b2mi::kEmuSharedBytes );
#endif
	{
		std::lock_guard<std::mutex> lock( g_stackMutex );
		for ( Fiber& f : L.fibers )
			g_stackPool.push_back( f.stack );
	}
	g_launch = outer;
}

// ---- runtime API ----

struct ihipEvent_t
{
	std::chrono::steady_clock::time_point t;
};

static ihipStream_t* const kStream = reinterpret_cast<ihipStream_t*>( 0x1 );

hipError_t hipGetDeviceCount( int* count )
{
	*count = 1;
	return hipSuccess;
}
hipError_t hipStreamDestroy( hipStream_t )
{
	return hipSuccess;
}
hipError_t hipStreamBeginCapture( hipStream_t, hipStreamCaptureMode )
{
	return hipErrorNotSupported;
}
hipError_t hipStreamEndCapture( hipStream_t, hipGraph_t* graph )
{
	*graph = nullptr;
	return hipErrorNotSupported;
}
hipError_t hipGraphInstantiate( hipGraphExec_t* exec, hipGraph_t, void*, char*, size_t )
{
	*exec = nullptr;
	return hipErrorNotSupported;
}
hipError_t hipGraphLaunch( hipGraphExec_t, hipStream_t )
{
	return hipErrorNotSupported;
}
hipError_t hipGraphExecDestroy( hipGraphExec_t )
{
	return hipSuccess;
}
hipError_t hipGraphDestroy( hipGraph_t )
{
	return hipSuccess;
}
hipError_t hipDeviceGetAttribute( int* value, hipDeviceAttribute_t attr, int )
{
	*value = attr == hipDeviceAttributeCooperativeLaunch ? 1 : attr == hipDeviceAttributeMultiprocessorCount ? 16 : 0;
	return hipSuccess;
}
hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor( int* blocks, const void*, int, size_t )
{
	*blocks = 4; // 16 x 4: at most 64 workgroups in one cooperative launch
	return hipSuccess;
}
hipError_t hipGetDevice( int* device )
{
	*device = 0;
	return hipSuccess;
}
hipError_t hipGetSymbolAddress( void** p, const void* symbol )
{
	*p = const_cast<void*>( symbol );
	return hipSuccess;
}
hipError_t hipMalloc( void** p, size_t bytes )
{
	*p = malloc( bytes ? bytes : 1 );
	if ( *p == nullptr )
		return hipErrorOutOfMemory;
	memset( *p, 0xff, bytes ); // device memory starts as garbage: NaN bytes
	return hipSuccess;
}
hipError_t hipFree( void* p )
{
	free( p );
	return hipSuccess;
}
hipError_t hipHostMalloc( void** p, size_t bytes, unsigned )
{
	*p = malloc( bytes ? bytes : 1 );
	return *p ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipHostFree( void* p )
{
	free( p );
	return hipSuccess;
}
hipError_t hipMemcpyAsync( void* dst, const void* src, size_t bytes, hipMemcpyKind, hipStream_t )
{
	if ( bytes )
		memcpy( dst, src, bytes );
	return hipSuccess;
}
hipError_t hipMemsetAsync( void* dst, int value, size_t bytes, hipStream_t )
{
	if ( bytes )
		memset( dst, value, bytes );
	return hipSuccess;
}
hipError_t hipStreamSynchronize( hipStream_t )
{
	return hipSuccess;
}
hipError_t hipStreamCreateWithFlags( hipStream_t* s, unsigned )
{
	*s = kStream;
	return hipSuccess;
}
hipError_t hipStreamWaitEvent( hipStream_t, hipEvent_t e, unsigned )
{
	return e ? hipSuccess : hipErrorInvalidValue;
}
hipError_t hipEventCreate( hipEvent_t* e )
{
	*e = new ihipEvent_t();
	return hipSuccess;
}
hipError_t hipEventCreateWithFlags( hipEvent_t* e, unsigned )
{
	return hipEventCreate( e );
}
hipError_t hipEventRecord( hipEvent_t e, hipStream_t )
{
	if ( e == nullptr )
		return hipErrorInvalidValue;
	e->t = std::chrono::steady_clock::now();
	return hipSuccess;
}
hipError_t hipEventElapsedTime( float* ms, hipEvent_t a, hipEvent_t b )
{
	if ( a == nullptr || b == nullptr )
		return hipErrorInvalidValue;
	*ms = std::chrono::duration<float, std::milli>( b->t - a->t ).count();
	return hipSuccess;
}
hipError_t hipEventDestroy( hipEvent_t e )
{
	delete e;
	return hipSuccess;
}
hipError_t hipFuncSetAttribute( const void*, hipFuncAttribute, int )
{
	return hipSuccess;
}
hipError_t hipGetLastError()
{
	return hipSuccess;
}
const char* hipGetErrorString( hipError_t e )
{
	return e == hipSuccess ? "no error" : "emulated HIP error";
}
