// Public API base macros for the MI355X step pipeline (box2c_amd).
//
// Drop-in re-declaration of the reference's public interface: the macro names and
// their meaning follow include/box2d/base.h of the reference so that unchanged
// application code compiles against this library.
#pragma once

#include <stdint.h>

#if defined( B2MI_BUILDING_LIBRARY )
	#define B2_EXPORT_ATTR __attribute__( ( visibility( "default" ) ) )
#else
	#define B2_EXPORT_ATTR
#endif

#if defined( __HIPCC__ )
	// the inline math is also compiled for the GPU: host and device share one definition
	#define B2_API extern "C" B2_EXPORT_ATTR
	#define B2_INLINE __host__ __device__ inline
	#define B2_LITERAL( T ) T
	#define B2_ZERO_INIT {}
#elif defined( __cplusplus )
	#define B2_API extern "C" B2_EXPORT_ATTR
	#define B2_INLINE inline
	#define B2_LITERAL( T ) T
	#define B2_ZERO_INIT {}
#else
	#define B2_API B2_EXPORT_ATTR
	#define B2_INLINE static inline
	#define B2_LITERAL( T ) ( T )
	#define B2_ZERO_INIT { 0 }
#endif

// Assertion hook (reference: include/box2d/base.h:35, src/core.c:23-37).
typedef int b2AssertFcn( const char* condition, const char* fileName, int lineNumber );
B2_API void b2SetAssertFcn( b2AssertFcn* assertFcn );

typedef struct b2Version
{
	int major;
	int minor;
	int revision;
} b2Version;

// reference: src/core.c:39-42 (returns 3.0.0)
B2_API b2Version b2GetVersion( void );
