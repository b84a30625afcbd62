// Reference-side tools, compiled together with the reference's own sources (never part of
// the product). Built by oracle/Makefile into oracle/_ref/libbox2d_ref.so.
//
//  1. b2mi_World_DumpContacts / b2mi_World_DumpBodies: the inspection records declared in
//     include/b2mi.h, read from the reference's internal structures (src/world.h,
//     src/constraint_graph.h, src/solver_set.h), the same way the reference's own tests
//     include src/ internals (test/CMakeLists.txt:22-23).
//  2. ref_pool_*: a pthread implementation of the reference's task plugin interface
//     (b2EnqueueTaskCallback / b2FinishTaskCallback, include/box2d/types.h:14-48), standing
//     in for enkiTS, which the reference fetches from the network (CMakeLists.txt:80-92).
//     Semantics follow benchmark/main.c:71-115: a task set is split into ranges of at least
//     minRange items, ranges run concurrently, the worker index passed to the callback is the
//     index of the executing thread (unique among concurrently running ranges), and finish
//     helps execute queued ranges while it waits.
#define _GNU_SOURCE
#include "body.h"
#include "constraint_graph.h"
#include "contact.h"
#include "core.h"
#include "solver_set.h"
#include "world.h"

#include "box2d/box2d.h"

#include <pthread.h>
#include <time.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>

// Record layouts must match include/b2mi.h of this repository (that header is not
// included here because it re-declares the public API that the reference headers declare).
typedef struct RefContactRecord
{
	int32_t contactId;
	int32_t shapeIdA;
	int32_t shapeIdB;
	int32_t colorIndex;
	int32_t localIndex;
	uint32_t simFlags;
	float friction;
	float restitution;
	b2Manifold manifold;
} RefContactRecord;

typedef struct RefBodyRecord
{
	int32_t bodyId;
	int32_t type;
	int32_t setIndex;
	int32_t localIndex;
	b2Transform transform;
	b2Vec2 center;
	b2Vec2 linearVelocity;
	float angularVelocity;
	float pad;
} RefBodyRecord;

static void ref_fill_contact( RefContactRecord* r, const b2ContactSim* cs, int color, int local )
{
	r->contactId = cs->contactId;
	r->shapeIdA = cs->shapeIdA;
	r->shapeIdB = cs->shapeIdB;
	r->colorIndex = color;
	r->localIndex = local;
	r->simFlags = cs->simFlags;
	r->friction = cs->friction;
	r->restitution = cs->restitution;
	r->manifold = cs->manifold;
}

int b2mi_World_DumpContacts( b2WorldId worldId, RefContactRecord* records, int capacity )
{
	b2World* world = b2GetWorldFromId( worldId );
	int n = 0;
	for ( int c = 0; c < b2_graphColorCount; ++c )
	{
		b2ContactArray* arr = &world->constraintGraph.colors[c].contacts;
		for ( int i = 0; i < arr->count; ++i, ++n )
		{
			if ( n < capacity )
				ref_fill_contact( records + n, arr->data + i, c, i );
		}
	}
	b2ContactArray* arr = &world->solverSetArray[b2_awakeSet].contacts;
	for ( int i = 0; i < arr->count; ++i, ++n )
	{
		if ( n < capacity )
			ref_fill_contact( records + n, arr->data + i, -1, i );
	}
	return n;
}

int b2mi_World_DumpBodies( b2WorldId worldId, RefBodyRecord* records, int capacity )
{
	b2World* world = b2GetWorldFromId( worldId );
	int count = b2Array( world->bodyArray ).count;
	int n = 0;
	for ( int id = 0; id < count; ++id )
	{
		b2Body* body = world->bodyArray + id;
		if ( body->setIndex == B2_NULL_INDEX )
			continue;
		if ( n < capacity )
		{
			RefBodyRecord* r = records + n;
			memset( r, 0, sizeof( *r ) );
			b2SolverSet* set = world->solverSetArray + body->setIndex;
			b2BodySim* sim = set->sims.data + body->localIndex;
			r->bodyId = id;
			r->type = body->type;
			r->setIndex = body->setIndex;
			r->localIndex = body->localIndex;
			r->transform = sim->transform;
			r->center = sim->center;
			if ( body->setIndex == b2_awakeSet )
			{
				b2BodyState* st = set->states.data + body->localIndex;
				r->linearVelocity = st->linearVelocity;
				r->angularVelocity = st->angularVelocity;
			}
		}
		n += 1;
	}
	return n;
}

// ---------------------------------------------------------------------------------------
// pthread task pool

#define POOL_MAX_THREADS 256
#define POOL_MAX_RANGES 4096
#define POOL_MAX_SETS 512

typedef struct PoolSet
{
	b2TaskCallback* fn;
	void* ctx;
	atomic_int remaining;
} PoolSet;

typedef struct PoolRange
{
	PoolSet* set;
	int start, end;
} PoolRange;

typedef struct Pool
{
	int threadCount;
	pthread_t threads[POOL_MAX_THREADS];
	pthread_mutex_t mutex;
	pthread_cond_t cond;
	PoolRange queue[POOL_MAX_RANGES];
	int head, tail; // ring buffer [head, tail)
	PoolSet sets[POOL_MAX_SETS];
	int setCursor;
	int quit;
} Pool;

static Pool g_pool;

typedef struct ThreadArg
{
	int index;
} ThreadArg;
static ThreadArg g_args[POOL_MAX_THREADS];

static int pool_pop_locked( PoolRange* out )
{
	if ( g_pool.head == g_pool.tail )
		return 0;
	*out = g_pool.queue[g_pool.head % POOL_MAX_RANGES];
	g_pool.head += 1;
	return 1;
}

static void pool_run( const PoolRange* r, int workerIndex )
{
	r->set->fn( r->start, r->end, (uint32_t)workerIndex, r->set->ctx );
	atomic_fetch_sub( &r->set->remaining, 1 );
}

static void* pool_thread( void* p )
{
	int index = ( (ThreadArg*)p )->index;
	for ( ;; )
	{
		pthread_mutex_lock( &g_pool.mutex );
		PoolRange r;
		while ( g_pool.quit == 0 && pool_pop_locked( &r ) == 0 )
			pthread_cond_wait( &g_pool.cond, &g_pool.mutex );
		if ( g_pool.quit )
		{
			pthread_mutex_unlock( &g_pool.mutex );
			return NULL;
		}
		pthread_mutex_unlock( &g_pool.mutex );
		pool_run( &r, index );
	}
}

// Starts threadCount-1 worker threads; the calling thread is worker 0.
void ref_pool_start( int threadCount )
{
	if ( threadCount < 1 )
		threadCount = 1;
	if ( threadCount > POOL_MAX_THREADS )
		threadCount = POOL_MAX_THREADS;
	memset( &g_pool, 0, sizeof( g_pool ) );
	pthread_mutex_init( &g_pool.mutex, NULL );
	pthread_cond_init( &g_pool.cond, NULL );
	g_pool.threadCount = threadCount;
	for ( int i = 1; i < threadCount; ++i )
	{
		g_args[i].index = i;
		pthread_create( g_pool.threads + i, NULL, pool_thread, g_args + i );
	}
}

void ref_pool_stop( void )
{
	pthread_mutex_lock( &g_pool.mutex );
	g_pool.quit = 1;
	pthread_cond_broadcast( &g_pool.cond );
	pthread_mutex_unlock( &g_pool.mutex );
	for ( int i = 1; i < g_pool.threadCount; ++i )
		pthread_join( g_pool.threads[i], NULL );
	g_pool.threadCount = 0;
}

static void* ref_pool_enqueue( b2TaskCallback* task, int itemCount, int minRange, void* taskContext, void* userContext )
{
	(void)userContext;
	if ( g_pool.threadCount <= 1 || itemCount <= minRange )
	{
		task( 0, itemCount, 0, taskContext );
		return NULL;
	}
	int maxRanges = g_pool.threadCount;
	int rangeCount = ( itemCount + minRange - 1 ) / minRange;
	if ( rangeCount > maxRanges )
		rangeCount = maxRanges;
	pthread_mutex_lock( &g_pool.mutex );
	PoolSet* set = g_pool.sets + ( g_pool.setCursor++ % POOL_MAX_SETS );
	set->fn = task;
	set->ctx = taskContext;
	atomic_store( &set->remaining, rangeCount );
	int per = itemCount / rangeCount, rem = itemCount % rangeCount, start = 0;
	for ( int i = 0; i < rangeCount; ++i )
	{
		int n = per + ( i < rem ? 1 : 0 );
		PoolRange r = { set, start, start + n };
		g_pool.queue[g_pool.tail % POOL_MAX_RANGES] = r;
		g_pool.tail += 1;
		start += n;
	}
	pthread_cond_broadcast( &g_pool.cond );
	pthread_mutex_unlock( &g_pool.mutex );
	return set;
}

static void ref_pool_finish( void* userTask, void* userContext )
{
	(void)userContext;
	PoolSet* set = userTask;
	if ( set == NULL )
		return;
	while ( atomic_load( &set->remaining ) > 0 )
	{
		PoolRange r;
		pthread_mutex_lock( &g_pool.mutex );
		int got = pool_pop_locked( &r );
		pthread_mutex_unlock( &g_pool.mutex );
		if ( got )
			pool_run( &r, 0 );
	}
}

// Solver tasks (one per worker) spin on each other, so each must get its own thread. A task
// set of itemCount 1 goes straight to the queue; with threadCount-1 pool threads plus the
// finishing main thread every solver worker runs concurrently.
static void* ref_pool_enqueue_any( b2TaskCallback* task, int itemCount, int minRange, void* taskContext, void* userContext )
{
	if ( g_pool.threadCount > 1 && itemCount == 1 && minRange == 1 )
	{
		pthread_mutex_lock( &g_pool.mutex );
		PoolSet* set = g_pool.sets + ( g_pool.setCursor++ % POOL_MAX_SETS );
		set->fn = task;
		set->ctx = taskContext;
		atomic_store( &set->remaining, 1 );
		PoolRange r = { set, 0, 1 };
		g_pool.queue[g_pool.tail % POOL_MAX_RANGES] = r;
		g_pool.tail += 1;
		pthread_cond_broadcast( &g_pool.cond );
		pthread_mutex_unlock( &g_pool.mutex );
		return set;
	}
	return ref_pool_enqueue( task, itemCount, minRange, taskContext, userContext );
}

// Points a world definition at the pool (workerCount = pool threads).
void ref_pool_apply( b2WorldDef* def )
{
	if ( g_pool.threadCount > 1 )
	{
		def->workerCount = g_pool.threadCount;
		def->enqueueTask = ref_pool_enqueue_any;
		def->finishTask = ref_pool_finish;
		def->userTaskContext = NULL;
	}
}

// ---- world-level parallelism: independent worlds stepped concurrently ----
//
// `threads` threads take worlds from a shared counter (pass them largest first for a good
// balance) and step each taken world `steps` times with the reference's default
// single-threaded task system. Worlds share no mutable state, so this is the reference's best
// CPU schedule for many small worlds.
typedef struct WorldsJob
{
	const b2WorldId* worlds;
	int count, steps, subSteps;
	float timeStep;
	atomic_int next;
} WorldsJob;

static void* worlds_thread( void* p )
{
	WorldsJob* job = (WorldsJob*)p;
	for ( ;; )
	{
		int i = atomic_fetch_add( &job->next, 1 );
		if ( i >= job->count )
			break;
		for ( int s = 0; s < job->steps; ++s )
			b2World_Step( job->worlds[i], job->timeStep, job->subSteps );
	}
	return NULL;
}

// Returns the wall-clock seconds of the whole run.
double ref_step_worlds_parallel( const b2WorldId* worlds, int count, int threads, int steps, float timeStep, int subSteps )
{
	if ( threads < 1 )
		threads = 1;
	if ( threads > POOL_MAX_THREADS )
		threads = POOL_MAX_THREADS;
	pthread_t tid[POOL_MAX_THREADS];
	WorldsJob job;
	job.worlds = worlds;
	job.count = count;
	job.steps = steps;
	job.subSteps = subSteps;
	job.timeStep = timeStep;
	atomic_init( &job.next, 0 );
	struct timespec t0, t1;
	clock_gettime( CLOCK_MONOTONIC, &t0 );
	for ( int t = 0; t < threads; ++t )
		pthread_create( tid + t, NULL, worlds_thread, &job );
	for ( int t = 0; t < threads; ++t )
		pthread_join( tid[t], NULL );
	clock_gettime( CLOCK_MONOTONIC, &t1 );
	return (double)( t1.tv_sec - t0.tv_sec ) + 1e-9 * (double)( t1.tv_nsec - t0.tv_nsec );
}
