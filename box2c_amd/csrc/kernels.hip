// HIP kernels of the step pipeline for CDNA4 (gfx950).
//
//  k_init_contacts  scatter new contacts created by the host pair update
//  k_collide        narrowphase: one thread per awake contact (src/world.c:290-365,
//                   src/contact.c:493-582, src/manifold.c:14-712)
//  k_solve          the whole solver of one step for one work item per workgroup:
//                   prepare, substeps of (integrate velocities, warm start, solve,
//                   integrate positions, relax) over the graph colours, restitution,
//                   store impulses and finalize (src/solver.c:34-351, 559-785,
//                   src/contact_solver.c:21-1293)
//
// Arithmetic: every expression keeps the reference's operation order and is compiled with
// -ffp-contract=off and correctly rounded division and square root, so each result is the
// same IEEE float the reference computes. min/max use the reference's x86 operand rule
// (a > b ? a : b), never v_max/v_min.
#include <hip/hip_runtime.h>

#include <float.h>
#include <type_traits>

#include "b2mi_layout.h"
#include "box2d/math_functions.h"

#include "narrowphase.hpp"
#include "joints.hpp"
#include "toi.hpp"

namespace b2mi
{


__device__ inline b2Transform sim_transform( const DevPose& s )
{
	b2Transform xf;
	xf.p = v2( s.p[0], s.p[1] );
	xf.q.c = s.q[0];
	xf.q.s = s.q[1];
	return xf;
}

__device__ inline bool aabb_overlaps( const float* a, const float* b )
{
	// src/aabb.h:59-71
	float d1x = b[0] - a[2], d1y = b[1] - a[3];
	float d2x = a[0] - b[2], d2y = a[1] - b[3];
	if ( d1x > 0.0f || d1y > 0.0f )
		return false;
	if ( d2x > 0.0f || d2y > 0.0f )
		return false;
	return true;
}

__global__ void k_init_contacts( const DevWorld* __restrict__ worlds, const ContactInit* __restrict__ inits, int count )
{
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if ( i >= count )
		return;
	ContactInit ci = inits[i];
	DevContact* c = worlds[ci.world].contacts + ci.contactId;
	DevContact z;
	memset( &z, 0, sizeof( z ) );
	z.shapeIdA = ci.shapeIdA;
	z.shapeIdB = ci.shapeIdB;
	z.friction = ci.friction;
	z.restitution = ci.restitution;
	*c = z;
	DevContactOut zo;
	memset( &zo, 0, sizeof( zo ) );
	zo.bodySimIndexA = -1;
	zo.bodySimIndexB = -1;
	worlds[ci.world].contactOut[ci.contactId] = zo;
}

// src/world.c:304-362 + src/contact.c:531-581 for one contact
// Workgroups are dealt to the 8 XCDs round-robin (block b runs on XCD b % 8). Remapping
// block b to logical block xcd_block(b) gives each XCD one contiguous eighth of the list, so
// the shapes and bodies a run of neighbouring contacts shares stay in that XCD's L2.
__device__ inline int xcd_block( int b, int nblocks )
{
	const int kXcds = 8;
	int q = nblocks / kXcds, r = nblocks % kXcds;
	int x = b % kXcds, pos = b / kXcds;
	return x * q + ( x < r ? x : r ) + pos;
}

// One contact of world W (index `worldIndex`): its shapes' ids come with the list entry, so their
// lines load beside the contact's own line rather than after it.
__device__ __forceinline__ void collide_contact( const DevWorld& W, int worldIndex, int contactId, int shapeIdA, int shapeIdB,
												 ChangeRecord* __restrict__ changes, int* __restrict__ changeCount )
{
	DevContact* c = W.contacts + contactId;
	const DevShape& A = W.shapes[shapeIdA];
	const DevShape& B = W.shapes[shapeIdB];
	uint32_t simFlags = c->simFlags;
	uint32_t change = 0;

	if ( aabb_overlaps( A.fatAABB, B.fatAABB ) == false )
	{
		simFlags |= kSimDisjoint;
		simFlags &= ~kSimTouching;
		change = kSimDisjoint;
	}
	else
	{
		bool wasTouching = ( simFlags & kSimTouching ) != 0;
		bool staticA = ( A.body & kStaticRef ) != 0, staticB = ( B.body & kStaticRef ) != 0;
		const DevPose sa = staticA ? W.staticPoses[A.body & ~kStaticRef] : W.poses[A.body];
		const DevPose sb = staticB ? W.staticPoses[B.body & ~kStaticRef] : W.poses[B.body];
		c->invMassA = sa.invMass;
		c->invIA = sa.invI;
		c->invMassB = sb.invMass;
		c->invIB = sb.invI;
		b2Transform xfA = sim_transform( sa ), xfB = sim_transform( sb );
		b2Vec2 offsetA = b2RotateVector( xfA.q, v2( sa.localCenter[0], sa.localCenter[1] ) );
		b2Vec2 offsetB = b2RotateVector( xfB.q, v2( sb.localCenter[0], sb.localCenter[1] ) );

		Manifold m;
		collide_shapes( m, A, xfA, B, xfB );
		bool touching = m.pointCount > 0;
		if ( touching && ( ( A.flags | B.flags ) & kShapeHitEvents ) )
			simFlags |= kSimHitEvent;
		else
			simFlags &= ~kSimHitEvent;

		// warm-start transfer by feature id (src/contact.c:555-568); the two points are
		// unrolled so the manifolds stay in registers
		uint16_t oldId[2] = { c->manifold.points[0].id, c->manifold.points[1].id };
		float oldNormal[2] = { c->manifold.points[0].normalImpulse, c->manifold.points[1].normalImpulse };
		float oldTangent[2] = { c->manifold.points[0].tangentImpulse, c->manifold.points[1].tangentImpulse };
		int oldCount = c->manifold.pointCount;
		DevManifold nm;
		memset( &nm, 0, sizeof( nm ) );
		uint32_t warmFrom = 0xFu;
		nm.normalX = m.normal.x;
		nm.normalY = m.normal.y;
		nm.pointCount = m.pointCount;
		DevContactOut out;
		out.bodySimIndexA = staticA ? -1 : A.body;
		out.bodySimIndexB = staticB ? -1 : B.body;
#pragma unroll
		for ( int k = 0; k < 2; ++k )
		{
			DevManifoldPoint& p = nm.points[k];
			out.pointX[k] = m.point[k].x;
			out.pointY[k] = m.point[k].y;
			out.normalVelocity[k] = 0.0f;
			p.anchorAX = m.anchorA[k].x;
			p.anchorAY = m.anchorA[k].y;
			p.anchorBX = m.anchorB[k].x;
			p.anchorBY = m.anchorB[k].y;
			p.separation = m.separation[k];
			p.id = m.id[k];
			if ( k < m.pointCount )
			{
				b2Vec2 aA = b2Sub( v2( p.anchorAX, p.anchorAY ), offsetA );
				b2Vec2 aB = b2Sub( v2( p.anchorBX, p.anchorBY ), offsetB );
				p.anchorAX = aA.x;
				p.anchorAY = aA.y;
				p.anchorBX = aB.x;
				p.anchorBY = aB.y;
				p.normalImpulse = 0.0f;
				p.tangentImpulse = 0.0f;
				p.maxNormalImpulse = 0.0f;
				p.persisted = 0;
				bool found = false;
#pragma unroll
				for ( int j = 0; j < 2; ++j )
				{
					if ( !found && j < oldCount && oldId[j] == p.id )
					{
						p.normalImpulse = oldNormal[j];
						p.tangentImpulse = oldTangent[j];
						p.persisted = 1;
						found = true;
						warmFrom = ( warmFrom & ~( 3u << ( 2 * k ) ) ) | ( (uint32_t)j << ( 2 * k ) );
					}
				}
			}
		}
		c->manifold = nm;
		c->warmFrom = warmFrom;
		W.contactOut[contactId] = out;
		if ( touching )
			simFlags |= kSimTouching;
		else
			simFlags &= ~kSimTouching;
		if ( touching && !wasTouching )
			change = kSimStarted;
		else if ( !touching && wasTouching )
			change = kSimStopped;
	}
	c->simFlags = simFlags;
	if ( change )
	{
		int slot = atomicAdd( changeCount, 1 );
		ChangeRecord r;
		r.world = worldIndex;
		r.contactId = contactId;
		r.flags = change;
		r.pad = 0;
		changes[slot] = r;
	}
}

__global__ void k_collide( const DevWorld* __restrict__ worlds, const int4* __restrict__ list, int count,
						   ChangeRecord* __restrict__ changes, int* __restrict__ changeCount )
{
	int i = xcd_block( blockIdx.x, gridDim.x ) * blockDim.x + threadIdx.x;
	if ( i >= count )
		return;
	const int4 item = list[i];
	if ( item.x < 0 )
		return; // unused entry of a world's region
	// a wave whose contacts all belong to one world (the usual case: each world's contacts are
	// one range of the list) reads that world's pointers with scalar loads
	// a wave whose contacts all belong to one world (the usual case: each world's contacts are
	// one range of the list) reads that world's pointers with scalar loads
	const int w0 = __builtin_amdgcn_readfirstlane( item.x );
	if ( __all( item.x == w0 ) )
		collide_contact( worlds[w0], w0, item.y, item.z, item.w, changes, changeCount );
	else
		collide_contact( worlds[item.x], item.x, item.y, item.z, item.w, changes, changeCount );
}

__device__ inline void shape_aabb( const DevShape& s, b2Transform xf, float* out )
{
	// src/shape.c:457-478 / src/geometry.c:376-429
	b2Vec2 lower, upper;
	if ( s.type == kCircle )
	{
		b2Vec2 p = b2TransformPoint( xf, v2( s.vn[0][0], s.vn[0][1] ) );
		float r = s.radius;
		lower = v2( p.x - r, p.y - r );
		upper = v2( p.x + r, p.y + r );
	}
	else if ( s.type == kCapsule || s.type == kSegment )
	{
		b2Vec2 a = b2TransformPoint( xf, v2( s.vn[0][0], s.vn[0][1] ) );
		b2Vec2 b = b2TransformPoint( xf, v2( s.vn[1][0], s.vn[1][1] ) );
		lower = b2Min( a, b );
		upper = b2Max( a, b );
		if ( s.type == kCapsule )
		{
			b2Vec2 r = v2( s.radius, s.radius );
			lower = b2Sub( lower, r );
			upper = b2Add( upper, r );
		}
	}
	else
	{
		lower = b2TransformPoint( xf, v2( s.vn[0][0], s.vn[0][1] ) );
		upper = lower;
		for ( int i = 1; i < s.count; ++i )
		{
			b2Vec2 v = b2TransformPoint( xf, v2( s.vert( i )[0], s.vert( i )[1] ) );
			lower = b2Min( lower, v );
			upper = b2Max( upper, v );
		}
		b2Vec2 r = v2( s.radius, s.radius );
		lower = b2Sub( lower, r );
		upper = b2Add( upper, r );
	}
	out[0] = lower.x;
	out[1] = lower.y;
	out[2] = upper.x;
	out[3] = upper.y;
}

__device__ inline bool aabb_contains( const float* a, const float* b )
{
	return a[0] <= b[0] && a[1] <= b[1] && b[2] <= a[2] && b[3] <= a[3];
}

// solver.c:165-351 for one body. enlCount[0] counts the enlarge records, enlCount[1] the bodies
// at the sleep time. A fast body keeps its start pose (center0, rotation0) and its shapes' AABBs
// for its sweep and leaves one kFastBody record: the continuous stage (k_continuous) finishes it.
__device__ void finalize_body( const DevWorld& W, int simIndex, float4 s0, float4 s1, int worldIndex, EnlargeRecord* enl,
							   int* enlCount, int enlCapacity )
{
	DevBodySim& sim = W.sims[simIndex];
	b2Vec2 v = v2( s0.x, s0.y );
	float w = s0.z;
	b2Vec2 center = b2Add( v2( sim.center[0], sim.center[1] ), v2( s1.x, s1.y ) );
	b2Rot dq;
	dq.c = s1.z;
	dq.s = s1.w;
	b2Rot q0;
	q0.c = sim.q[0];
	q0.s = sim.q[1];
	b2Rot q = b2NormalizeRot( b2MulRot( dq, q0 ) );
	float maxVelocity = b2Length( v ) + b2AbsFloat( w ) * sim.maxExtent;
	float maxDeltaPosition = b2Length( v2( s1.x, s1.y ) ) + b2AbsFloat( dq.s ) * sim.maxExtent;
	float positionSleepFactor = 0.5f;
	float sleepVelocity = b2MaxFloat( maxVelocity, positionSleepFactor * W.inv_dt * maxDeltaPosition );
	b2Vec2 localCenter = v2( sim.localCenter[0], sim.localCenter[1] );
	b2Vec2 p = b2Sub( center, b2RotateVector( q, localCenter ) );

	sim.center[0] = center.x;
	sim.center[1] = center.y;
	sim.q[0] = q.c;
	sim.q[1] = q.s;
	sim.p[0] = p.x;
	sim.p[1] = p.y;
	DevPose pose;
	pose.p[0] = p.x;
	pose.p[1] = p.y;
	pose.q[0] = q.c;
	pose.q[1] = q.s;
	pose.localCenter[0] = localCenter.x;
	pose.localCenter[1] = localCenter.y;
	pose.invMass = sim.invMass;
	pose.invI = sim.invI;
	W.poses[simIndex] = pose;
	sim.force[0] = 0.0f;
	sim.force[1] = 0.0f;
	sim.torque = 0.0f;

	bool isFast = false;
	bool enableSleep = ( W.flags & kWorldSleep ) != 0 && ( sim.flags & kBodyEnableSleep ) != 0;
	if ( enableSleep == false || sleepVelocity > sim.sleepThreshold )
	{
		sim.sleepTime = 0.0f;
		const float safetyFactor = 0.5f;
		if ( ( sim.flags & kBodyDynamic ) && ( W.flags & kWorldContinuous ) && maxVelocity * W.dt > safetyFactor * sim.minExtent )
			isFast = true;
	}
	else
	{
		// falling asleep (src/solver.c:271-278)
		sim.sleepTime += W.dt;
	}
	// a body at the sleep time lets its island sleep when the rest of it is there too
	// (src/solver.c:281-298); the host checks the islands of such bodies
	if ( sim.sleepTime >= K_TIME_TO_SLEEP )
		atomicAdd( enlCount + 1, 1 );
	if ( isFast )
	{
		int slot = atomicAdd( enlCount, 1 );
		if ( slot < enlCapacity )
		{
			EnlargeRecord r;
			r.world = worldIndex;
			r.simIndex = simIndex;
			r.shapeId = -1;
			r.flags = kFastBody | ( ( sim.flags & kBodyBullet ) ? kBulletFast : 0u );
			for ( int c = 0; c < 4; ++c )
				r.fatAABB[c] = 0.0f;
			enl[slot] = r;
		}
		return;
	}
	sim.center0[0] = center.x;
	sim.center0[1] = center.y;
	sim.rotation0[0] = q.c;
	sim.rotation0[1] = q.s;

	b2Transform xf;
	xf.p = p;
	xf.q = q;
	for ( int k = 0; k < sim.shapeCount; ++k )
	{
		int shapeId = W.bodyShapes[sim.shapeStart + k];
		DevShape& shape = W.shapes[shapeId];
		float aabb[4];
		shape_aabb( shape, xf, aabb );
		aabb[0] -= K_SPECULATIVE;
		aabb[1] -= K_SPECULATIVE;
		aabb[2] += K_SPECULATIVE;
		aabb[3] += K_SPECULATIVE;
		for ( int c = 0; c < 4; ++c )
			shape.aabb[c] = aabb[c];
		if ( aabb_contains( shape.fatAABB, aabb ) == false )
		{
			shape.fatAABB[0] = aabb[0] - K_AABB_MARGIN;
			shape.fatAABB[1] = aabb[1] - K_AABB_MARGIN;
			shape.fatAABB[2] = aabb[2] + K_AABB_MARGIN;
			shape.fatAABB[3] = aabb[3] + K_AABB_MARGIN;
			int slot = atomicAdd( enlCount, 1 );
			if ( slot < enlCapacity )
			{
				EnlargeRecord r;
				r.world = worldIndex;
				r.simIndex = simIndex;
				r.shapeId = shapeId;
				r.flags = kEnlarged;
				for ( int c = 0; c < 4; ++c )
					r.fatAABB[c] = shape.fatAABB[c];
				enl[slot] = r;
			}
		}
	}
}

// This thread's first slot of the next non-empty colour group after g (cyclically over the
// groups 1..kGroups-1), or -1: the colour sweeps visit the same slots in every pass, so a store
// can load a slot's read-only quads one phase ahead.
__device__ inline int next_colour_slot( const WorkItem& wi, int g, int tid )
{
	int n = next_group( wi, g );
	if ( n == 0 )
		return -1;
	int b = wi.groupStart[n], e = wi.groupStart[n + 1];
	return b + tid < e ? b + tid : -1;
}

// How the threads that run one work item synchronize between dependent phases.
//
// BlockSync: one workgroup, workgroup barriers, the restitution vote in LDS.
struct BlockSync
{
	static constexpr bool kGrid = false;
	int levelRun; // overflow levels of an HBM-store item: the largest cached run (slots)
	__device__ inline int tid() const
	{
		return threadIdx.x;
	}
	__device__ inline int nt() const
	{
		return blockDim.x;
	}
	__device__ inline void sync() const
	{
		__syncthreads();
	}
};

// GridSync: every workgroup of a cooperative launch (all resident at once), a grid-wide barrier
// between phases and the vote in global memory. The barrier counts arrivals on one word that the
// host zeroes before the launch: barrier n waits until n x blocks workgroups have arrived. The
// arrival is a device-scope release and the wait a device-scope acquire, after a workgroup
// barrier: each phase's writes by any thread (body state, impulses, in whichever XCD's L2 they
// are) are visible to every workgroup of the next phase. Every
// wait is bounded: past kGridSpinLimit polls it sets the error word, which ends every other wait
// at once and which the host turns into an error after the step.
constexpr unsigned kGridSpinLimit = 1u << 24;
struct GridSync
{
	static constexpr bool kGrid = true;
	unsigned* arrivals;
	int* vote;
	int* error;
	int tid_, nt_;
	unsigned blocks;
	int levelRun; // overflow levels of an HBM-store item: the largest cached run (slots), run by workgroup 0
	mutable unsigned target;
	__device__ inline int tid() const
	{
		return tid_;
	}
	__device__ inline int nt() const
	{
		return nt_;
	}
	__device__ inline void sync() const
	{
		__syncthreads();
		if ( threadIdx.x == 0 )
		{
			target += blocks;
			__hip_atomic_fetch_add( arrivals, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT );
			unsigned polls = 0;
			while ( __hip_atomic_load( arrivals, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT ) < target )
			{
				if ( __hip_atomic_load( error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT ) != 0 )
					break;
				__builtin_amdgcn_s_sleep( 1 );
				if ( ++polls > kGridSpinLimit )
				{
					__hip_atomic_store( error, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT );
					break;
				}
			}
		}
		__syncthreads();
	}
};

// One run of overflow levels [L, E) of lv (slot bounds within the item, at most kLevelCacheSlots
// slots in all) of an HBM-store item, for one overflow contact or joint level phase (`phase`, a
// SplitPhase), by the threads of one workgroup, all of which call it. It stages in LDS the
// distinct bodies the run's slots touch (state and masses, found through an LDS hash table), each
// slot's two bodies as cache indices, and for contacts each slot's record and point count; runs
// the levels on that copy with a workgroup barrier after each level; then writes back what the
// levels write (the bodies' velocities and the records' written quads) and ends with a workgroup
// barrier. The same functions on the same values in the same level order as on the store itself,
// so the same bits. Kept out of line: the solver kernels call it only between colour sweeps, and
// inlined its registers would add to theirs.
__device__ __attribute__( ( noinline ) ) void cached_level_run( const GlobalStore src, const DevWorld& W, const int* __restrict__ sc,
															   const int* __restrict__ lv, int L, int E, int phase )
{
	__shared__ RunCache rc;
	const int tid = threadIdx.x, nt = blockDim.x;
	const int s0 = lv[L], s1 = lv[E];
	if ( s1 - s0 > kLevelCacheSlots )
		return; // never: the callers form runs of at most kLevelCacheSlots slots
	const bool contacts = phase == kSplitOverflowWarmStart || phase == kSplitOverflowSolve || phase == kSplitOverflowRelax ||
						  phase == kSplitOverflowRestitution;
	constexpr unsigned mask = kRunCacheHash - 1;
	for ( int h = tid; h < kRunCacheHash; h += nt )
		rc.hashKey[h] = -1;
	if ( tid == 0 )
		rc.count = 0;
	__syncthreads();
	// the distinct dynamic bodies (at most two per slot, so the table is at most half full)
	for ( int k = s0 + tid; k < s1; k += nt )
	{
		const int2 b = src.bodies( k );
#pragma unroll
		for ( int j = 0; j < 2; ++j )
		{
			const int x = j == 0 ? b.x : b.y;
			if ( x < 0 )
				continue;
			unsigned h = ( (unsigned)x * 2654435761u ) & mask;
			for ( ;; )
			{
				const int prev = atomicCAS( &rc.hashKey[h], -1, x );
				if ( prev == -1 || prev == x )
					break;
				h = ( h + 1 ) & mask;
			}
		}
	}
	__syncthreads();
	for ( int h = tid; h < kRunCacheHash; h += nt )
	{
		const int x = rc.hashKey[h];
		if ( x >= 0 )
		{
			const int i = atomicAdd( &rc.count, 1 );
			rc.hashVal[h] = i;
			rc.key[i] = x;
		}
	}
	__syncthreads();
	const int n = rc.count;
	for ( int i = tid; i < n; i += nt )
	{
		const int b = rc.key[i];
		rc.P[2 * i] = src.b0( b );
		rc.P[2 * i + 1] = src.b1( b );
		rc.M[i] = src.mass( b );
	}
	auto find = [&]( int x ) {
		if ( x < 0 )
			return x;
		unsigned h = ( (unsigned)x * 2654435761u ) & mask;
		while ( rc.hashKey[h] != x )
			h = ( h + 1 ) & mask;
		return rc.hashVal[h];
	};
	for ( int k = s0 + tid; k < s1; k += nt )
	{
		const int2 b = src.bodies( k );
		rc.sb[k - s0] = make_int2( find( b.x ), find( b.y ) );
		rc.pc[k - s0] = contacts ? W.contacts[sc[k]].manifold.pointCount : 0;
		if ( contacts )
		{
#pragma unroll
			for ( int q = 0; q < cfQuads; ++q )
				rc.rec[( k - s0 ) * cfQuads + q] = src.quad( k, q );
		}
	}
	__syncthreads();
	const RunStore C{ &rc, s0 };
	const Softness soft = world_softness( W );
	for ( int M = L; M < E; ++M )
	{
		for ( int k = lv[M] + tid; k < lv[M + 1]; k += nt )
		{
			const int2 b = rc.sb[k - s0];
			const int pc = rc.pc[k - s0];
			switch ( phase )
			{
				case kSplitOverflowWarmStart:
					ovf_warm_start( C, k, b.x, b.y, pc );
					break;
				case kSplitOverflowSolve:
				case kSplitOverflowRelax:
					ovf_solve( C, k, b.x, b.y, pc, W.inv_h, W.pushout, phase == kSplitOverflowSolve, soft );
					break;
				case kSplitOverflowRestitution:
					ovf_restitution( C, k, b.x, b.y, pc, W.restitutionThreshold );
					break;
				case kSplitJointWarmStart:
					warm_start_joint( C, W.joints[sc[k]], b.x, b.y );
					break;
				case kSplitJointSolve:
				case kSplitJointRelax:
					solve_joint( C, W.joints[sc[k]], b.x, b.y, W, phase == kSplitJointSolve );
					break;
				default:
					break;
			}
		}
		__syncthreads();
	}
	for ( int i = tid; i < n; i += nt )
		src.b0( rc.key[i] ) = rc.P[2 * i];
	if ( contacts )
		for ( int k = s0 + tid; k < s1; k += nt )
		{
#pragma unroll
			for ( int q = cfMutableQuad; q < cfQuads; ++q )
				src.setQuad( k, q, rc.rec[( k - s0 ) * cfQuads + q] );
		}
	__syncthreads();
}

// All stages of one step for one work item, by the threads of X (one workgroup, or a cooperative
// grid for a very large item), with X's barriers between dependent phases. `C` holds the item's
// private state; `bl` its body list (sim indices), `sc` the contact or joint id of each of its
// slots and `sb` each slot's two local bodies (in HBM; the joints' are read from there).
template <class S, class Sync>
__device__ void solve_item( const S& C, const Sync& X, const DevWorld& W, const WorkItem& wi, const int* __restrict__ bl,
							const int* __restrict__ sc, const int2* __restrict__ sb, const int* __restrict__ levels,
							const int* __restrict__ jointLevels, EnlargeRecord* enl, int* enlCount, int enlCapacity )
{
	const int tid = X.tid(), nt = X.nt();
	const float4* gstate = reinterpret_cast<const float4*>( W.state );
	const int total = wi.groupStart[kGroups];
	// the threads' vote on skipping the colour restitution sweeps (set here, visible after the
	// staging barrier; cleared by any thread that finds a slot or body that needs the sweeps)
	__shared__ int blockVote;
	if ( tid == 0 )
	{
		if constexpr ( Sync::kGrid )
			__hip_atomic_store( X.vote, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT );
		else
			blockVote = 1;
	}

	// stage private body state and masses (and the static body's identity state)
	if constexpr ( S::kIdentityBody )
		if ( tid == 0 )
		{
			C.b0( C.identity ) = make_float4( 0.0f, 0.0f, 0.0f, 0.0f );
			C.b1( C.identity ) = make_float4( 0.0f, 0.0f, 1.0f, 0.0f );
			C.mass( C.identity ) = make_float2( 0.0f, 0.0f );
		}
	for ( int i = tid; i < wi.bodyCount; i += nt )
	{
		int b = bl[i];
		C.b0( i ) = gstate[2 * b];
		C.b1( i ) = gstate[2 * b + 1];
		C.mass( i ) = make_float2( W.sims[b].invMass, W.sims[b].invI );
	}
	X.sync();

	// prepare the joints (src/solver.c:598-606: before the contacts; they write only their own
	// records) and the contacts (all groups; the overflow group uses the same per-point formulas),
	// with the contact of a thread's next slot loaded while it prepares the current one
	const int jointBegin = wi.jointGroupStart[0], jointEnd = wi.jointGroupStart[kGroups];
	for ( int k = jointBegin + tid; k < jointEnd; k += nt )
	{
		int2 lb = sb[k];
		prepare_joint( W.joints[sc[k]], joint_body_sim( W, bl, lb.x ), joint_body_sim( W, bl, lb.y ), W );
	}
	{
		PrepInput cur;
		if ( tid < total )
			cur = load_prep( W.contacts + sc[tid] );
		for ( int k = tid; k < total; k += nt )
		{
			PrepInput next;
			if ( k + nt < total )
				next = load_prep( W.contacts + sc[k + nt] );
			int2 lb = C.bodies( k );
			prepare_slot( C, k, cur, lb.x, lb.y, W );
			cur = next;
		}
	}
	X.sync();

	// read-only quads of this thread's slot in the next colour group (q0 holds its bodies),
	// loaded one phase ahead, before the barrier: the sweeps visit the same slots in every pass,
	// so after the barrier a lane starts on its bodies at once
	float4 pre[cfMutableQuad];
	int preSlot = -1;
	// (the slot, nk, is looked up at the start of a phase and its quads loaded at the end)
	auto ahead = [&]( int nk ) {
		if constexpr ( S::kPrefetch )
		{
			if ( nk >= 0 )
				C.loadQuads( nk, pre );
			preSlot = nk;
		}
	};
	const int ovfBegin = wi.groupStart[0], ovfEnd = wi.groupStart[1];
	// the overflow group runs level by level (slots of one level share no changed body)
	const int levelCount = ovfEnd > ovfBegin ? wi.ovfLevelCount : 0;
	const float inv_h = W.inv_h, h = W.h;
	const float pushout = W.pushout, minBiasVel = -W.pushout;
	const bool warm = ( W.flags & kWorldWarmStart ) != 0;
	const Softness soft = world_softness( W );

	// the overflow joints run level by level before the overflow contacts, the joints of a colour
	// beside its contacts (a colour's joints and contacts share no body), src/solver.c:638-681
	const int jointLevelCount = wi.jointGroupStart[1] > wi.jointGroupStart[0] ? wi.jointLevelCount : 0;
	auto jointStage = [&]( int b, int e, int stage ) {
		for ( int k = b + tid; k < e; k += nt )
		{
			int2 lb = sb[k];
			if ( stage == 0 )
				warm_start_joint( C, W.joints[sc[k]], lb.x, lb.y );
			else
				solve_joint( C, W.joints[sc[k]], lb.x, lb.y, W, stage == 1 );
		}
	};
	// Runs levels [0, count) of lv in order, the slots of a level in parallel (they share no body
	// that any of them changes): `body( store, slot, bodyA, bodyB, pointCount )`. An HBM-store item
	// runs consecutive levels of at most X.levelRun slots in all as one cached run
	// (cached_level_run; in a grid by workgroup 0 alone, then one grid barrier); a larger level,
	// and every level of the LDS stores, runs over all of X's threads with X's barrier after it.
	auto runLevels = [&]( const int* lv, int count, int phase, auto&& body ) {
		const bool contacts = phase < kSplitJointPrepare;
		auto plain = [&]( int L ) {
			for ( int k = lv[L] + tid; k < lv[L + 1]; k += nt )
			{
				const int2 lb = contacts ? C.bodies( k ) : sb[k];
				body( k, lb.x, lb.y, contacts ? W.contacts[sc[k]].manifold.pointCount : 0 );
			}
			X.sync();
		};
		if constexpr ( std::is_same<S, GlobalStore>::value )
		{
			const int run = X.levelRun < kLevelCacheSlots ? X.levelRun : kLevelCacheSlots;
			int L = 0;
			while ( L < count )
			{
				if ( run <= 0 || lv[L + 1] - lv[L] > run )
				{
					plain( L );
					++L;
					continue;
				}
				int E = L + 1;
				while ( E < count && lv[E + 1] - lv[L] <= run )
					++E;
				if constexpr ( Sync::kGrid )
				{
					if ( blockIdx.x == 0 )
						cached_level_run( C, W, sc, lv, L, E, phase );
					X.sync();
				}
				else
					cached_level_run( C, W, sc, lv, L, E, phase );
				L = E;
			}
		}
		else
			for ( int L = 0; L < count; ++L )
				plain( L );
	};
	auto overflowJoints = [&]( int stage ) {
		runLevels( jointLevels, jointLevelCount, stage == 0 ? kSplitJointWarmStart : stage == 1 ? kSplitJointSolve : kSplitJointRelax,
				   [&]( int k, int la, int lb, int ) {
					   if ( stage == 0 )
						   warm_start_joint( C, W.joints[sc[k]], la, lb );
					   else
						   solve_joint( C, W.joints[sc[k]], la, lb, W, stage == 1 );
				   } );
	};

	// the velocity terms of this thread's first two bodies stay in registers over the substeps
	VelocityTerms vt[2] = {};
#pragma unroll
	for ( int r = 0; r < 2; ++r )
		if ( tid + r * nt < wi.bodyCount )
			vt[r] = velocity_terms( W.sims[bl[tid + r * nt]], W );

	for ( int sub = 0; sub < W.subSteps; ++sub )
	{
#pragma unroll
		for ( int r = 0; r < 2; ++r )
			if ( tid + r * nt < wi.bodyCount )
				integrate_velocity_t( C, tid + r * nt, vt[r], W );
		for ( int i = tid + 2 * nt; i < wi.bodyCount; i += nt )
			integrate_velocity( C, i, W.sims[bl[i]], W );
		X.sync();

		// warm start: overflow (always, src/solver.c:639), then colours if enabled
		overflowJoints( 0 );
		runLevels( levels, levelCount, kSplitOverflowWarmStart, [&]( int k, int la, int lb, int pc ) {
			ovf_warm_start( C, k, la, lb, pc );
		} );
		if ( warm )
			for ( int g = 1; g < kGroups; ++g )
			{
				int b = wi.groupStart[g], e = wi.groupStart[g + 1];
				int jb = wi.jointGroupStart[g], je = wi.jointGroupStart[g + 1];
				if ( b == e && jb == je )
					continue;
				jointStage( jb, je, 0 );
				if ( b < e )
				{
					const int nk = S::kPrefetch ? next_colour_slot( wi, g, tid ) : -1;
					for ( int k = b + tid; k < e; k += nt )
					{
						if ( S::kPrefetch && k == preSlot )
						{
							int2 lb = quad_bodies( pre[0] );
							lane_warm_start_q( C, k, lb.x, lb.y, pre );
						}
						else
						{
							int2 lb = C.bodies( k );
							lane_warm_start( C, k, lb.x, lb.y );
						}
					}
					ahead( nk );
				}
				X.sync();
			}

		// solve with bias
		overflowJoints( 1 );
		runLevels( levels, levelCount, kSplitOverflowSolve, [&]( int k, int la, int lb, int pc ) {
			ovf_solve( C, k, la, lb, pc, inv_h, pushout, true, soft );
		} );
		for ( int g = 1; g < kGroups; ++g )
		{
			int b = wi.groupStart[g], e = wi.groupStart[g + 1];
			int jb = wi.jointGroupStart[g], je = wi.jointGroupStart[g + 1];
			if ( b == e && jb == je )
				continue;
			jointStage( jb, je, 1 );
			if ( b < e )
			{
				const int nk = S::kPrefetch ? next_colour_slot( wi, g, tid ) : -1;
				for ( int k = b + tid; k < e; k += nt )
				{
					if ( S::kPrefetch && k == preSlot )
					{
						int2 lb = quad_bodies( pre[0] );
						lane_solve_q( C, k, lb.x, lb.y, pre, inv_h, minBiasVel, true, soft );
					}
					else
					{
						int2 lb = C.bodies( k );
						lane_solve( C, k, lb.x, lb.y, inv_h, minBiasVel, true, soft );
					}
				}
				ahead( nk );
			}
			X.sync();
		}

		for ( int i = tid; i < wi.bodyCount; i += nt )
			integrate_position( C, i, h );
		X.sync();

		// relax
		overflowJoints( 2 );
		runLevels( levels, levelCount, kSplitOverflowRelax, [&]( int k, int la, int lb, int pc ) {
			ovf_solve( C, k, la, lb, pc, inv_h, pushout, false, soft );
		} );
		for ( int g = 1; g < kGroups; ++g )
		{
			int b = wi.groupStart[g], e = wi.groupStart[g + 1];
			int jb = wi.jointGroupStart[g], je = wi.jointGroupStart[g + 1];
			if ( b == e && jb == je )
				continue;
			jointStage( jb, je, 2 );
			if ( b < e )
			{
				const int nk = S::kPrefetch ? next_colour_slot( wi, g, tid ) : -1;
				for ( int k = b + tid; k < e; k += nt )
				{
					if ( S::kPrefetch && k == preSlot )
					{
						int2 lb = quad_bodies( pre[0] );
						lane_solve_q( C, k, lb.x, lb.y, pre, inv_h, minBiasVel, false, soft );
					}
					else
					{
						int2 lb = C.bodies( k );
						lane_solve( C, k, lb.x, lb.y, inv_h, minBiasVel, false, soft );
					}
				}
				ahead( nk );
			}
			X.sync();
		}
	}

	// restitution
	const float threshold = W.restitutionThreshold;
	runLevels( levels, levelCount, kSplitOverflowRestitution, [&]( int k, int la, int lb, int pc ) {
		ovf_restitution( C, k, la, lb, pc, threshold );
	} );
	// The colour sweeps would return their input bit for bit when every colour slot's stage is
	// the identity and every body is plain (restitution_is_identity): then they are skipped.
	bool identity = W.subSteps > 0;
	for ( int k = ovfEnd + tid; k < total && identity; k += nt )
		identity = slot_restitution_is_identity( C, k, threshold );
	for ( int i = tid; i < wi.bodyCount && identity; i += nt )
		identity = body_is_plain( C, i );
	if ( !identity )
	{
		// every writer writes the same value
		if constexpr ( Sync::kGrid )
			__hip_atomic_store( X.vote, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT );
		else
			blockVote = 0;
	}
	X.sync();
	if constexpr ( Sync::kGrid )
		identity = __hip_atomic_load( X.vote, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT ) != 0;
	else
		identity = blockVote != 0;
	for ( int g = 1; g < kGroups && !identity; ++g )
	{
		int b = wi.groupStart[g], e = wi.groupStart[g + 1];
		if ( b == e )
			continue;
		const int nk = S::kPrefetch ? next_colour_slot( wi, g, tid ) : -1;
		for ( int k = b + tid; k < e; k += nt )
		{
			if ( S::kPrefetch && k == preSlot )
			{
				int2 lb = quad_bodies( pre[0] );
				lane_restitution_q( C, k, lb.x, lb.y, pre, threshold );
			}
			else
			{
				int2 lb = C.bodies( k );
				lane_restitution( C, k, lb.x, lb.y, threshold );
			}
		}
		ahead( nk );
		X.sync();
	}

	// store impulses (contact_solver.c:442-471 overflow: pointCount points; 1180-1293 lanes: both);
	// a lane loads its next slot's contact id while it stores the current one
	int contactNext = tid < total ? sc[tid] : 0;
	for ( int k = tid; k < total; k += nt )
	{
		const int contact = contactNext;
		if ( k + nt < total )
			contactNext = sc[k + nt];
		DevManifold& m = W.contacts[contact].manifold;
		DevContactOut& out = W.contactOut[contact];
		int n = k < ovfEnd ? m.pointCount : 2;
#pragma unroll
		for ( int j = 0; j < 2; ++j )
		{
			if ( j >= n )
				break;
			float4 imp = C.quad( k, cfMutableQuad + j );
			m.points[j].normalImpulse = imp.x;
			m.points[j].tangentImpulse = imp.y;
			m.points[j].maxNormalImpulse = imp.z;
			out.normalVelocity[j] = imp.w;
		}
	}

	// finalize owned bodies and write their state back (velocity kept, deltas reset)
	float4* gs = reinterpret_cast<float4*>( W.state );
	int bodyNext = tid < wi.ownedCount ? bl[tid] : 0;
	for ( int i = tid; i < wi.ownedCount; i += nt )
	{
		const int b = bodyNext;
		if ( i + nt < wi.ownedCount )
			bodyNext = bl[i + nt];
		float4 s0 = C.b0( i ), s1 = C.b1( i );
		finalize_body( W, b, s0, s1, wi.world, enl, enlCount, enlCapacity );
		gs[2 * b] = s0;
		gs[2 * b + 1] = make_float4( 0.0f, 0.0f, 1.0f, 0.0f );
	}
}

// the dynamic LDS block of the solver kernels (its size is given at launch)
extern __shared__ float4 lds[];

// One workgroup per work item; one kernel per store (every item of a launch has that store).
// kStoreLds items keep their state in the dynamic LDS block (layout: body state, constraint
// records, masses), kStoreBodiesLds items the body state and the written quads, kStoreHbm items
// nothing.
template <int Store>
__global__ void __launch_bounds__( kSolveMaxBlock ) k_solve( const DevWorld* __restrict__ worlds, const WorkItem* __restrict__ items,
												   const int* __restrict__ bodyList, const int* __restrict__ slotContact,
												   const int2* __restrict__ slotBodies, const int* __restrict__ ovfLevels, float* cfBase,
												   float4* priv, float2* mass, EnlargeRecord* enl, int* enlCount, int enlCapacity,
												   int itemOffset, int levelRun, const int* abortFlag )
{
	if ( abortFlag != nullptr && *abortFlag != 0 )
		return;
	// read in place (uniform scalar loads); a private copy would be indexed dynamically from scratch
	const WorkItem& wi = items[itemOffset + blockIdx.x];
	const DevWorld& W = worlds[wi.world];
	const int* bl = bodyList + wi.bodyStart;
	const int* sc = slotContact + wi.slotStart;
	const int total = wi.groupStart[kGroups];
	if constexpr ( Store == kStoreLds )
	{
		// LDS (float4 arrays first, 16-byte aligned): body state, constraint records, masses
		LdsStore C;
		C.P = lds;
		C.identity = wi.bodyCount;
		C.rec = lds + 2 * ( wi.bodyCount + 1 );
		C.M = reinterpret_cast<float2*>( C.rec + cfQuads * total );
		for ( int k = threadIdx.x; k < total; k += blockDim.x )
		{
			int2 b = slotBodies[wi.slotStart + k];
			C.rec[k * cfQuads] = make_float4( __builtin_bit_cast( float, b.x ), __builtin_bit_cast( float, b.y ), 0.0f, 0.0f );
		}
		// the staging loop inside solve_item ends with a barrier before any C.bodies() read
		solve_item( C, BlockSync{ levelRun }, W, wi, bl, sc, slotBodies + wi.slotStart, ovfLevels + wi.ovfLevelBegin,
					ovfLevels + wi.jointLevelBegin, enl, enlCount, enlCapacity );
	}
	else if constexpr ( Store == kStoreBodiesLds )
	{
		// LDS: body state, written quads, masses, slot bodies; HBM: the read-only quads, loaded one
		// colour phase ahead
		BodiesLdsStore C;
		C.rec = reinterpret_cast<float4*>( cfBase );
		C.base = wi.slotStart;
		C.P = lds;
		C.identity = wi.bodyCount;
		C.mut = lds + 2 * ( wi.bodyCount + 1 );
		C.M = reinterpret_cast<float2*>( C.mut + 2 * total );
		C.sb = reinterpret_cast<int2*>( C.M + wi.bodyCount + 1 );
		for ( int k = threadIdx.x; k < total; k += blockDim.x )
			C.sb[k] = slotBodies[wi.slotStart + k];
		solve_item( C, BlockSync{ levelRun }, W, wi, bl, sc, slotBodies + wi.slotStart, ovfLevels + wi.ovfLevelBegin,
					ovfLevels + wi.jointLevelBegin, enl, enlCount, enlCapacity );
	}
	else
	{
		GlobalStore C;
		C.rec = reinterpret_cast<float4*>( cfBase );
		C.base = wi.slotStart;
		C.sb = slotBodies + wi.slotStart;
		C.P = priv + 2 * (size_t)wi.bodyStart;
		C.M = mass + wi.bodyStart;
		solve_item( C, BlockSync{ levelRun }, W, wi, bl, sc, slotBodies + wi.slotStart, ovfLevels + wi.ovfLevelBegin,
					ovfLevels + wi.jointLevelBegin, enl, enlCount, enlCapacity );
	}
}


// One very large HBM-store work item over a cooperative grid of workgroups (all resident at once):
// solve_item with grid-wide barriers where the workgroup path has workgroup barriers. The phases
// apply the same per-slot and per-body functions to the same data (a colour's slots share no
// body, so which thread takes which slot does not matter; overflow levels as in the workgroup
// path), so the results are the same bits. sync: [0] barrier arrivals (zero at launch), [1] the
// restitution vote; error: set when a barrier wait exceeded its bound; abortFlag: as k_solve's.
__global__ void __launch_bounds__( kGridSolveBlock ) k_solve_grid( const DevWorld* __restrict__ worlds, const WorkItem* __restrict__ items,
																 int itemIndex, const int* __restrict__ bodyList,
																 const int* __restrict__ slotContact, const int2* __restrict__ slotBodies,
																 const int* __restrict__ ovfLevels, float* cfBase, float4* priv, float2* mass,
																 EnlargeRecord* enl, int* enlCount, int enlCapacity, unsigned* sync, int* error,
																 int levelRun, const int* abortFlag )
{
	// (the flag was written before the launch began, so every workgroup reads the same value)
	if ( abortFlag != nullptr && *abortFlag != 0 )
		return;
	const WorkItem& wi = items[itemIndex];
	const DevWorld& W = worlds[wi.world];
	GlobalStore C;
	C.rec = reinterpret_cast<float4*>( cfBase );
	C.base = wi.slotStart;
	C.sb = slotBodies + wi.slotStart;
	C.P = priv + 2 * (size_t)wi.bodyStart;
	C.M = mass + wi.bodyStart;
	GridSync X;
	X.arrivals = sync;
	X.vote = reinterpret_cast<int*>( sync + 1 );
	X.error = error;
	X.tid_ = blockIdx.x * blockDim.x + threadIdx.x;
	X.nt_ = gridDim.x * blockDim.x;
	X.blocks = gridDim.x;
	X.levelRun = levelRun;
	X.target = 0;
	solve_item( C, X, W, wi, bodyList + wi.bodyStart, slotContact + wi.slotStart, slotBodies + wi.slotStart,
				ovfLevels + wi.ovfLevelBegin, ovfLevels + wi.jointLevelBegin, enl, enlCount, enlCapacity );
}

__global__ void k_grid_reset( unsigned* sync )
{
	if ( threadIdx.x < 2 )
		sync[threadIdx.x] = 0u;
}

// One slot or body of one phase of a very large work item (HBM store): `k` is the body or slot
// index within the item. The host launches the phases in solve_item's order, so every workgroup
// barrier of solve_item becomes a kernel boundary; each phase applies the same per-body and
// per-slot functions to the same data, so the results are the same bits (a colour's slots share
// no body, so their order does not matter; overflow levels as in solve_item).
__device__ void split_work( int phase, int k, const WorkItem& wi, const DevWorld& W, const GlobalStore& C, const int* __restrict__ bl,
							const int* __restrict__ sc, EnlargeRecord* enl, int* enlCount, int enlCapacity )
{
	const Softness soft = world_softness( W );
	const float minBiasVel = -W.pushout;
	switch ( phase )
	{
		case kSplitStage:
		{
			const float4* gstate = reinterpret_cast<const float4*>( W.state );
			int b = bl[k];
			C.b0( k ) = gstate[2 * b];
			C.b1( k ) = gstate[2 * b + 1];
			C.mass( k ) = make_float2( W.sims[b].invMass, W.sims[b].invI );
			break;
		}
		case kSplitPrepare:
		{
			int2 lb = C.bodies( k );
			prepare_slot( C, k, load_prep( W.contacts + sc[k] ), lb.x, lb.y, W );
			break;
		}
		case kSplitJointPrepare:
		{
			int2 lb = C.bodies( k );
			prepare_joint( W.joints[sc[k]], joint_body_sim( W, bl, lb.x ), joint_body_sim( W, bl, lb.y ), W );
			break;
		}
		case kSplitJointWarmStart:
		{
			int2 lb = C.bodies( k );
			warm_start_joint( C, W.joints[sc[k]], lb.x, lb.y );
			break;
		}
		case kSplitJointSolve:
		case kSplitJointRelax:
		{
			int2 lb = C.bodies( k );
			solve_joint( C, W.joints[sc[k]], lb.x, lb.y, W, phase == kSplitJointSolve );
			break;
		}
		case kSplitIntegrateVelocities:
			integrate_velocity( C, k, W.sims[bl[k]], W );
			break;
		case kSplitIntegratePositions:
			integrate_position( C, k, W.h );
			break;
		case kSplitOverflowWarmStart:
		{
			int2 lb = C.bodies( k );
			ovf_warm_start( C, k, lb.x, lb.y, W.contacts[sc[k]].manifold.pointCount );
			break;
		}
		case kSplitWarmStart:
		{
			int2 lb = C.bodies( k );
			lane_warm_start( C, k, lb.x, lb.y );
			break;
		}
		case kSplitOverflowSolve:
		case kSplitOverflowRelax:
		{
			int2 lb = C.bodies( k );
			ovf_solve( C, k, lb.x, lb.y, W.contacts[sc[k]].manifold.pointCount, W.inv_h, W.pushout, phase == kSplitOverflowSolve,
					   soft );
			break;
		}
		case kSplitSolve:
		case kSplitRelax:
		{
			int2 lb = C.bodies( k );
			lane_solve( C, k, lb.x, lb.y, W.inv_h, minBiasVel, phase == kSplitSolve, soft );
			break;
		}
		case kSplitOverflowRestitution:
		{
			int2 lb = C.bodies( k );
			ovf_restitution( C, k, lb.x, lb.y, W.contacts[sc[k]].manifold.pointCount, W.restitutionThreshold );
			break;
		}
		case kSplitRestitution:
		{
			int2 lb = C.bodies( k );
			lane_restitution( C, k, lb.x, lb.y, W.restitutionThreshold );
			break;
		}
		case kSplitStoreImpulses:
		{
			const int contact = sc[k];
			DevManifold& m = W.contacts[contact].manifold;
			DevContactOut& out = W.contactOut[contact];
			int n = k < wi.groupStart[1] ? m.pointCount : 2;
#pragma unroll
			for ( int j = 0; j < 2; ++j )
			{
				if ( j >= n )
					break;
				float4 imp = C.quad( k, cfMutableQuad + j );
				m.points[j].normalImpulse = imp.x;
				m.points[j].tangentImpulse = imp.y;
				m.points[j].maxNormalImpulse = imp.z;
				out.normalVelocity[j] = imp.w;
			}
			break;
		}
		case kSplitFinalize:
		{
			float4* gs = reinterpret_cast<float4*>( W.state );
			int b = bl[k];
			float4 s0 = C.b0( k ), s1 = C.b1( k );
			finalize_body( W, b, s0, s1, wi.world, enl, enlCount, enlCapacity );
			gs[2 * b] = s0;
			gs[2 * b + 1] = make_float4( 0.0f, 0.0f, 1.0f, 0.0f );
			break;
		}
		default:
			break;
	}
}

__device__ inline GlobalStore split_store( const WorkItem& wi, const int2* slotBodies, float* cfBase, float4* priv, float2* mass )
{
	GlobalStore C;
	C.rec = reinterpret_cast<float4*>( cfBase );
	C.base = wi.slotStart;
	C.sb = slotBodies + wi.slotStart;
	C.P = priv + 2 * (size_t)wi.bodyStart;
	C.M = mass + wi.bodyStart;
	return C;
}

// One phase of one very large work item spread over a grid: thread i < jointCount takes joint
// slot `jointFirst + i` of jointPhase, the next `count` threads slot or body `first + i` of phase
// (a colour's joints and contacts share no body, so they run side by side).
__global__ void __launch_bounds__( 256 ) k_split_phase( const DevWorld* __restrict__ worlds, const WorkItem* __restrict__ items,
														int itemIndex, int phase, int first, int count, int jointPhase, int jointFirst,
														int jointCount, const int* __restrict__ bodyList, const int* __restrict__ slotContact,
														const int2* __restrict__ slotBodies, float* cfBase, float4* priv, float2* mass,
														EnlargeRecord* enl, int* enlCount, int enlCapacity )
{
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if ( i >= jointCount + count )
		return;
	const WorkItem& wi = items[itemIndex];
	const DevWorld& W = worlds[wi.world];
	const GlobalStore C = split_store( wi, slotBodies, cfBase, priv, mass );
	const int* bl = bodyList + wi.bodyStart;
	const int* sc = slotContact + wi.slotStart;
	if ( i < jointCount )
		split_work( jointPhase, jointFirst + i, wi, W, C, bl, sc, enl, enlCount, enlCapacity );
	else
		split_work( phase, first + i - jointCount, wi, W, C, bl, sc, enl, enlCount, enlCapacity );
}

// One cached run of overflow levels of one very large work item in one workgroup: levels
// [levelBegin, levelBegin + levelCount) of `levels` (slot bounds within the item, at most
// kLevelCacheSlots slots in all), an overflow contact or joint level phase (cached_level_run).
__global__ void __launch_bounds__( kSplitLevelBlock ) k_split_levels( const DevWorld* __restrict__ worlds, const WorkItem* __restrict__ items,
																	 int itemIndex, int phase, const int* __restrict__ levels, int levelBegin,
																	 int levelCount, const int* __restrict__ slotContact,
																	 const int2* __restrict__ slotBodies, float* cfBase, float4* priv, float2* mass )
{
	const WorkItem& wi = items[itemIndex];
	cached_level_run( split_store( wi, slotBodies, cfBase, priv, mass ), worlds[wi.world], slotContact + wi.slotStart, levels + levelBegin, 0,
					  levelCount, phase );
}

// ---- continuous collision (src/solver.c:787-1066) ----

// b2MakeShapeDistanceProxy (src/shape.c:713-734): every shape keeps its points in vert(i)
__device__ inline b2DistanceProxy shape_proxy( const DevShape& s )
{
	b2Vec2 pts[b2_maxPolygonVertices];
	const int n = s.count < b2_maxPolygonVertices ? s.count : b2_maxPolygonVertices;
	for ( int i = 0; i < n; ++i )
		pts[i] = v2( s.vert( i )[0], s.vert( i )[1] );
	return make_proxy( pts, n, s.type == kSegment ? 0.0f : s.radius );
}

// b2GetShapeCentroid (src/shape.c:480-497)
__device__ inline b2Vec2 shape_centroid( const DevShape& s )
{
	if ( s.type == kPolygon )
		return v2( s.centroid[0], s.centroid[1] );
	if ( s.type == kCircle )
		return v2( s.vert( 0 )[0], s.vert( 0 )[1] );
	return b2Lerp( v2( s.vert( 0 )[0], s.vert( 0 )[1] ), v2( s.vert( 1 )[0], s.vert( 1 )[1] ), 0.5f );
}

// b2MakeSweep (src/body.h:159-168) of an awake body, and of a static body, whose center is its
// origin (src/body.c:530-532) and which does not move
__device__ inline b2Sweep sim_sweep( const DevBodySim& sim )
{
	b2Sweep s;
	s.localCenter = v2( sim.localCenter[0], sim.localCenter[1] );
	s.c1 = v2( sim.center0[0], sim.center0[1] );
	s.c2 = v2( sim.center[0], sim.center[1] );
	s.q1.c = sim.rotation0[0];
	s.q1.s = sim.rotation0[1];
	s.q2.c = sim.q[0];
	s.q2.s = sim.q[1];
	return s;
}

__device__ inline b2Sweep static_sweep( const DevPose& pose )
{
	b2Sweep s;
	s.localCenter = v2( 0.0f, 0.0f );
	s.c1 = s.c2 = v2( pose.p[0], pose.p[1] );
	s.q1.c = s.q2.c = pose.q[0];
	s.q1.s = s.q2.s = pose.q[1];
	return s;
}

// b2ShouldShapesCollide (src/shape.h:75-83)
__device__ inline bool filters_collide( const DevShapeFilter& a, const DevShapeFilter& b )
{
	if ( a.groupIndex == b.groupIndex && a.groupIndex != 0 )
		return a.groupIndex > 0;
	return ( a.maskBits & b.categoryBits ) != 0 && ( a.categoryBits & b.maskBits ) != 0;
}

constexpr int kContinuousStack = 256;

// src/solver.c:904-1066 for fast body i of `list` (one thread each): each shape's sweep from
// its last AABB to its final one is queried against the static tree (and, for a bullet, the
// kinematic and dynamic trees) in the reference's depth-first order, and the time of impact
// against every candidate that passes the reference's filters (src/solver.c:796-901) keeps the
// smallest fraction. With a hit the body moves back to that fraction of its sweep; either way
// its start pose becomes its final one and its shapes' fat AABBs grow where they must
// (kFastEnlarged records). hits counts the bodies moved back; error is set when a query stack
// overflows (the step then fails).
__global__ void k_continuous( const DevWorld* __restrict__ worlds, const ContinuousBody* __restrict__ list, int count,
							  const int* __restrict__ noCollide, EnlargeRecord* enl, int* enlCount, int enlCapacity, int* hits, int* error )
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if ( i >= count )
		return;
	const ContinuousBody& cb = list[i];
	const DevWorld& W = worlds[cb.world];
	DevBodySim& sim = W.sims[cb.simIndex];
	const b2Sweep sweep = sim_sweep( sim );
	const b2Transform xf2 = { b2Sub( sweep.c2, b2RotateVector( sweep.q2, sweep.localCenter ) ), sweep.q2 };
	const bool bullet = ( sim.flags & kBodyBullet ) != 0;
	float fraction = 1.0f;
	int stack[kContinuousStack];

	for ( int k = 0; k < sim.shapeCount; ++k )
	{
		const int fastId = W.bodyShapes[sim.shapeStart + k];
		DevShape& fs = W.shapes[fastId];
		float box2[4];
		shape_aabb( fs, xf2, box2 );
		const float box[4] = { minr( fs.aabb[0], box2[0] ), minr( fs.aabb[1], box2[1] ), maxr( fs.aabb[2], box2[2] ),
							   maxr( fs.aabb[3], box2[3] ) };
		for ( int c = 0; c < 4; ++c )
			fs.aabb[c] = box2[c];
		const DevShapeFilter fastFilter = cb.filters[fastId];
		const b2DistanceProxy fastProxy = shape_proxy( fs );

		// the candidate shape `shapeId` of the query (b2ContinuousQueryCallback)
		auto candidate = [&]( int shapeId ) {
			if ( shapeId == fastId )
				return;
			const DevShape& other = W.shapes[shapeId];
			if ( other.body == fs.body )
				return;
			if ( filters_collide( fastFilter, cb.filters[shapeId] ) == false )
				return;
			b2Sweep otherSweep;
			if ( other.body & kStaticRef )
				otherSweep = static_sweep( W.staticPoses[other.body & ~kStaticRef] );
			else
			{
				const DevBodySim& os = W.sims[other.body];
				if ( os.flags & kBodyBullet )
					return;
				otherSweep = sim_sweep( os );
			}
			for ( int j = cb.noCollideBegin; j < cb.noCollideEnd; ++j )
				if ( noCollide[j] == other.body )
					return;
			b2TOIInput input;
			input.proxyA = shape_proxy( other );
			input.proxyB = fastProxy;
			input.sweepA = otherSweep;
			input.sweepB = sweep;
			input.tMax = fraction;
			b2TOIOutput out = time_of_impact( input );
			if ( 0.0f < out.t && out.t < fraction )
				fraction = out.t;
			else if ( 0.0f == out.t )
			{
				// a small circle around the fast shape's centroid
				const b2Vec2 centroid = shape_centroid( fs );
				input.proxyB = make_proxy( &centroid, 1, K_SPECULATIVE );
				out = time_of_impact( input );
				if ( 0.0f < out.t && out.t < fraction )
					fraction = out.t;
			}
		};

		const int trees = bullet ? 3 : 1;
		for ( int t = 0; t < trees; ++t )
		{
			const DevTreeNode* nodes = cb.nodes[t];
			if ( nodes == nullptr || cb.roots[t] < 0 )
				continue;
			// b2DynamicTree_Query (src/dynamic_tree.c:1120-1159): child2's subtree first
			int top = 0;
			stack[top++] = cb.roots[t];
			while ( top > 0 )
			{
				const int id = stack[--top];
				const DevTreeNode node = nodes[id];
				if ( node.categoryBits == 0u || box[0] > node.upperX || box[1] > node.upperY || node.lowerX > box[2] ||
					 node.lowerY > box[3] )
					continue;
				if ( node.child1 == -1 )
				{
					candidate( node.userData );
					continue;
				}
				if ( top + 2 > kContinuousStack )
				{
					atomicExch( error, 1 );
					return;
				}
				stack[top++] = node.child1;
				stack[top++] = node.child2;
			}
		}
	}

	b2Transform xf = xf2;
	float pad = 0.0f;
	if ( fraction < 1.0f )
	{
		// back to the time of impact
		const b2Rot q = b2NLerp( sweep.q1, sweep.q2, fraction );
		const b2Vec2 c = b2Lerp( sweep.c1, sweep.c2, fraction );
		xf.q = q;
		xf.p = b2Sub( c, b2RotateVector( q, sweep.localCenter ) );
		sim.p[0] = xf.p.x;
		sim.p[1] = xf.p.y;
		sim.q[0] = q.c;
		sim.q[1] = q.s;
		sim.center[0] = c.x;
		sim.center[1] = c.y;
		DevPose& pose = W.poses[cb.simIndex];
		pose.p[0] = xf.p.x;
		pose.p[1] = xf.p.y;
		pose.q[0] = q.c;
		pose.q[1] = q.s;
		pad = K_SPECULATIVE;
		atomicAdd( hits, 1 );
	}
	sim.center0[0] = sim.center[0];
	sim.center0[1] = sim.center[1];
	sim.rotation0[0] = sim.q[0];
	sim.rotation0[1] = sim.q[1];
	for ( int k = 0; k < sim.shapeCount; ++k )
	{
		const int shapeId = W.bodyShapes[sim.shapeStart + k];
		DevShape& shape = W.shapes[shapeId];
		if ( fraction < 1.0f )
		{
			float aabb[4];
			shape_aabb( shape, xf, aabb );
			shape.aabb[0] = aabb[0] - pad;
			shape.aabb[1] = aabb[1] - pad;
			shape.aabb[2] = aabb[2] + pad;
			shape.aabb[3] = aabb[3] + pad;
		}
		if ( aabb_contains( shape.fatAABB, shape.aabb ) == false )
		{
			shape.fatAABB[0] = shape.aabb[0] - K_AABB_MARGIN;
			shape.fatAABB[1] = shape.aabb[1] - K_AABB_MARGIN;
			shape.fatAABB[2] = shape.aabb[2] + K_AABB_MARGIN;
			shape.fatAABB[3] = shape.aabb[3] + K_AABB_MARGIN;
			int slot = atomicAdd( enlCount, 1 );
			if ( slot < enlCapacity )
			{
				EnlargeRecord r;
				r.world = cb.world;
				r.simIndex = cb.simIndex;
				r.shapeId = shapeId;
				r.flags = kFastEnlarged | ( bullet ? kBulletFast : 0u );
				for ( int c = 0; c < 4; ++c )
					r.fatAABB[c] = shape.fatAABB[c];
				enl[slot] = r;
			}
		}
	}
}

} // namespace b2mi

// ---- host launchers ----
namespace b2mi
{

hipError_t launch_init_contacts( const DevWorld* worlds, const ContactInit* inits, int count, hipStream_t stream )
{
	if ( count <= 0 )
		return hipSuccess;
	int block = 256, grid = ( count + block - 1 ) / block;
	hipLaunchKernelGGL( k_init_contacts, dim3( grid ), dim3( block ), 0, stream, worlds, inits, count );
	return hipGetLastError();
}

hipError_t launch_collide( const DevWorld* worlds, const int4* list, int count, ChangeRecord* changes, int* changeCount,
						   hipStream_t stream )
{
	if ( count <= 0 )
		return hipSuccess;
	int block = 256, grid = ( count + block - 1 ) / block;
	hipLaunchKernelGGL( k_collide, dim3( grid ), dim3( block ), 0, stream, worlds, list, count, changes, changeCount );
	return hipGetLastError();
}

hipError_t launch_solve( int store, const DevWorld* worlds, const WorkItem* items, int itemOffset, int itemCount, int blockSize,
						 int ldsBytes, const int* bodyList, const int* slotContact, const int2* slotBodies, const int* ovfLevels,
						 float* cf, float4* priv, float2* mass, EnlargeRecord* enl, int* enlCount, int enlCapacity,
						 int levelRun, const int* abortFlag, hipStream_t stream )
{
	if ( itemCount <= 0 )
		return hipSuccess;
	if ( ldsBytes < 0 || ldsBytes > kMaxItemLdsBytes || blockSize < 64 || blockSize > kSolveMaxBlock )
		return hipErrorInvalidValue;
	// HBM-store items use no dynamic LDS (the kernel's static LDS, the overflow run cache, would
	// not leave room for the largest dynamic block)
	if ( store == kStoreHbm && ldsBytes != 0 )
		return hipErrorInvalidValue;
	const void* kernels[3] = { reinterpret_cast<const void*>( &k_solve<kStoreHbm> ),
							   reinterpret_cast<const void*>( &k_solve<kStoreBodiesLds> ),
							   reinterpret_cast<const void*>( &k_solve<kStoreLds> ) };
	if ( store < 0 || store > 2 )
		return hipErrorInvalidValue;
	static int allowedLds[3] = { 0, 0, 0 }; // dynamic LDS above 64 KiB must be enabled per kernel
	if ( ldsBytes > 64 * 1024 && allowedLds[store] < kMaxItemLdsBytes )
	{
		hipError_t e = hipFuncSetAttribute( kernels[store], hipFuncAttributeMaxDynamicSharedMemorySize, kMaxItemLdsBytes );
		if ( e != hipSuccess )
			return e;
		allowedLds[store] = kMaxItemLdsBytes;
	}
#define B2MI_LAUNCH_SOLVE( S )                                                                                                  \
	hipLaunchKernelGGL( k_solve<S>, dim3( itemCount ), dim3( blockSize ), ldsBytes, stream, worlds, items, bodyList, slotContact, \
						slotBodies, ovfLevels, cf, priv, mass, enl, enlCount, enlCapacity, itemOffset, levelRun, abortFlag )
	if ( store == kStoreLds )
		B2MI_LAUNCH_SOLVE( kStoreLds );
	else if ( store == kStoreBodiesLds )
		B2MI_LAUNCH_SOLVE( kStoreBodiesLds );
	else
		B2MI_LAUNCH_SOLVE( kStoreHbm );
#undef B2MI_LAUNCH_SOLVE
	return hipGetLastError();
}


int solve_grid_capacity()
{
	static int capacity = -1;
	if ( capacity < 0 )
	{
		int device = 0, cooperative = 0, cus = 0, perCu = 0;
		capacity = 0;
		if ( hipGetDevice( &device ) == hipSuccess &&
			 hipDeviceGetAttribute( &cooperative, hipDeviceAttributeCooperativeLaunch, device ) == hipSuccess && cooperative &&
			 hipDeviceGetAttribute( &cus, hipDeviceAttributeMultiprocessorCount, device ) == hipSuccess &&
			 hipOccupancyMaxActiveBlocksPerMultiprocessor( &perCu, reinterpret_cast<const void*>( &k_solve_grid ), kGridSolveBlock, 0 ) ==
				 hipSuccess )
			capacity = cus * perCu;
	}
	return capacity;
}

hipError_t launch_solve_grid( const DevWorld* worlds, const WorkItem* items, int itemIndex, int blocks, const int* bodyList,
							  const int* slotContact, const int2* slotBodies, const int* ovfLevels, float* cf, float4* priv, float2* mass,
							  EnlargeRecord* enl, int* enlCount, int enlCapacity, unsigned* sync, int* error, int levelRun,
							  const int* abortFlag, hipStream_t stream )
{
	if ( blocks < 1 || blocks > solve_grid_capacity() )
		return hipErrorInvalidValue;
	hipLaunchKernelGGL( k_grid_reset, dim3( 1 ), dim3( 64 ), 0, stream, sync );
	hipError_t e = hipGetLastError();
	if ( e != hipSuccess )
		return e;
	void* args[] = { &worlds, &items, &itemIndex, &bodyList, &slotContact, &slotBodies, &ovfLevels, &cf,
					 &priv, &mass, &enl, &enlCount, &enlCapacity, &sync, &error, &levelRun, &abortFlag };
	return hipLaunchCooperativeKernel( k_solve_grid, dim3( blocks ), dim3( kGridSolveBlock ), args, 0, stream );
}

hipError_t launch_split_phase( const DevWorld* worlds, const WorkItem* items, int itemIndex, int phase, int first, int count,
							   int jointPhase, int jointFirst, int jointCount, const int* bodyList, const int* slotContact,
							   const int2* slotBodies, float* cf, float4* priv, float2* mass, EnlargeRecord* enl, int* enlCount,
							   int enlCapacity, hipStream_t stream )
{
	count = count > 0 ? count : 0;
	jointCount = jointCount > 0 ? jointCount : 0;
	if ( count + jointCount == 0 )
		return hipSuccess;
	if ( ( count > 0 && ( phase < 0 || phase >= kSplitPhaseCount || first < 0 ) ) ||
		 ( jointCount > 0 && ( jointPhase < kSplitJointPrepare || jointPhase > kSplitJointRelax || jointFirst < 0 ) ) )
		return hipErrorInvalidValue;
	int block = 256, grid = ( count + jointCount + block - 1 ) / block;
	hipLaunchKernelGGL( k_split_phase, dim3( grid ), dim3( block ), 0, stream, worlds, items, itemIndex, phase, first, count, jointPhase,
						jointFirst, jointCount, bodyList, slotContact, slotBodies, cf, priv, mass, enl, enlCount, enlCapacity );
	return hipGetLastError();
}

hipError_t launch_continuous( const DevWorld* worlds, const ContinuousBody* list, int count, const int* noCollide, EnlargeRecord* enl,
							  int* enlCount, int enlCapacity, int* hits, int* error, hipStream_t stream )
{
	if ( count <= 0 )
		return hipSuccess;
	int block = 64, grid = ( count + block - 1 ) / block;
	hipLaunchKernelGGL( k_continuous, dim3( grid ), dim3( block ), 0, stream, worlds, list, count, noCollide, enl, enlCount, enlCapacity,
						hits, error );
	return hipGetLastError();
}

hipError_t launch_split_levels( const DevWorld* worlds, const WorkItem* items, int itemIndex, int phase, const int* levels, int levelBegin,
								int levelCount, const int* bodyList, const int* slotContact, const int2* slotBodies, float* cf,
								float4* priv, float2* mass, hipStream_t stream )
{
	if ( levelCount <= 0 )
		return hipSuccess;
	// the level phases only (they touch no enlarge record)
	const bool level = phase == kSplitOverflowWarmStart || phase == kSplitOverflowSolve || phase == kSplitOverflowRelax ||
					   phase == kSplitOverflowRestitution || phase == kSplitJointWarmStart || phase == kSplitJointSolve ||
					   phase == kSplitJointRelax;
	if ( !level || levelBegin < 0 )
		return hipErrorInvalidValue;
	hipLaunchKernelGGL( k_split_levels, dim3( 1 ), dim3( kSplitLevelBlock ), 0, stream, worlds, items, itemIndex, phase, levels, levelBegin,
						levelCount, slotContact, slotBodies, cf, priv, mass );
	return hipGetLastError();
}

} // namespace b2mi
