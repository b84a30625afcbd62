// Solver arithmetic of one step, shared by the step kernel (kernels.hip) and the replay and
// adjoint kernels (adjoint.hip): the stores that hold a work item's state, body loads,
// prepare, the lane (graph colour) and overflow stages, and the two integrators. Each function
// keeps the reference's operation order (src/contact_solver.c, src/solver.c); with
// -ffp-contract=off and correctly rounded division and square root the results are the
// reference's floats.
#pragma once

#include "narrowphase.hpp"

namespace b2mi
{

// ------------------------------------------------------------------------------------
// solver

struct Body
{
	float vx, vy, w, flags;
	float dpx, dpy, dqc, dqs;
};

template <class S>
__device__ inline Body load_body( const S& C, int i )
{
	Body b;
	if constexpr ( S::kIdentityBody )
	{
		// the static body's identity state is stored at index C.identity: no branch
		float4 s0 = C.b0( i < 0 ? C.identity : i ), s1 = C.b1( i < 0 ? C.identity : i );
		b.vx = s0.x;
		b.vy = s0.y;
		b.w = s0.z;
		b.flags = s0.w;
		b.dpx = s1.x;
		b.dpy = s1.y;
		b.dqc = s1.z;
		b.dqs = s1.w;
		return b;
	}
	if ( i < 0 )
	{
		// identity state of a static body (src/body.h:90)
		b.vx = 0.0f;
		b.vy = 0.0f;
		b.w = 0.0f;
		b.flags = 0.0f;
		b.dpx = 0.0f;
		b.dpy = 0.0f;
		b.dqc = 1.0f;
		b.dqs = 0.0f;
		return b;
	}
	float4 s0 = C.b0( i ), s1 = C.b1( i );
	b.vx = s0.x;
	b.vy = s0.y;
	b.w = s0.z;
	b.flags = s0.w;
	b.dpx = s1.x;
	b.dpy = s1.y;
	b.dqc = s1.z;
	b.dqs = s1.w;
	return b;
}

template <class S>
__device__ inline void store_velocity( const S& C, int i, const Body& b )
{
	if ( i < 0 )
		return;
	C.b0( i ) = make_float4( b.vx, b.vy, b.w, b.flags );
}

// Solver state of one work item, addressed by local slot k (0 .. slots-1) and local body i:
// each slot's constraint record (seven float4, see ConstraintField), each slot's two local body
// indices, the private body state (two float4 per body: v.x, v.y, w, flags | dp.x, dp.y, dq.c,
// dq.s) and each body's inverse mass and inertia (the same floats the narrowphase copies into a
// contact). Every store offers the record as fields (at) and as quads (quad, setQuad).
//
// GlobalStore keeps it in HBM (any work item size). LdsStore keeps it in the workgroup's LDS,
// so the substep passes (about 3 x colours x substeps dependent phases per step) run at LDS
// latency. BodiesLdsStore, for items too large for that, keeps the bodies, the slot bodies and
// the written quads in LDS and reads the read-only quads from HBM, coalesced across lanes. The
// host picks the store per work item from the LDS budget. The stores that read the quads from HBM
// have the colour sweeps load a slot's read-only quads (loadQuads) one phase ahead (kPrefetch).
struct GlobalStore
{
	float4* rec; // record of slot k at rec[(base + k) * cfQuads]
	int base;
	const int2* sb;
	float4* P;
	float2* M;
	static constexpr bool kIdentityBody = false;
	__device__ inline float& at( int field, int k ) const
	{
		return reinterpret_cast<float*>( rec + (size_t)( base + k ) * cfQuads )[field];
	}
	__device__ inline float4 quad( int k, int q ) const
	{
		return rec[(size_t)( base + k ) * cfQuads + q];
	}
	__device__ inline void setQuad( int k, int q, float4 v ) const
	{
		rec[(size_t)( base + k ) * cfQuads + q] = v;
	}
	static constexpr bool kPrefetch = true;
	__device__ inline void loadQuads( int k, float4 ( &r )[cfMutableQuad] ) const
	{
#pragma unroll
		for ( int q = 0; q < cfMutableQuad; ++q )
			r[q] = quad( k, q );
	}
	__device__ inline int2 bodies( int k ) const
	{
		return sb[k];
	}
	__device__ inline float4& b0( int i ) const
	{
		return P[2 * i];
	}
	__device__ inline float4& b1( int i ) const
	{
		return P[2 * i + 1];
	}
	__device__ inline float2& mass( int i ) const
	{
		return M[i];
	}
};

struct BodiesLdsStore
{
	// HBM: the read-only quads of slot k at rec[base * cfQuads + k * cfMutableQuad], an 80-byte
	// stride inside the item's own record region (which starts at quad base * cfQuads, as an HBM
	// store item's does, so items of different stores never share records)
	float4* rec;
	int base;
	// the colour sweeps load a slot's read-only quads one phase ahead (loadQuads)
	static constexpr bool kPrefetch = true;
	float4* mut; // LDS: the written quads of slot k at mut[2k], mut[2k + 1]
	int2* sb;
	float4* P;
	float2* M;
	// the static body's identity state (v = 0, dp = 0, dq = (1, 0), no mass) at this index
	static constexpr bool kIdentityBody = true;
	int identity;
	__device__ inline float& at( int field, int k ) const
	{
		if ( field >= 4 * cfMutableQuad )
			return reinterpret_cast<float*>( mut + 2 * k )[field - 4 * cfMutableQuad];
		return reinterpret_cast<float*>( rec + (size_t)base * cfQuads + (size_t)k * cfMutableQuad )[field];
	}
	__device__ inline float4 quad( int k, int q ) const
	{
		if ( q >= cfMutableQuad )
			return mut[2 * k + q - cfMutableQuad];
		return rec[(size_t)base * cfQuads + (size_t)k * cfMutableQuad + q];
	}
	__device__ inline void loadQuads( int k, float4 ( &r )[cfMutableQuad] ) const
	{
		const float4* p = rec + (size_t)base * cfQuads + (size_t)k * cfMutableQuad;
#pragma unroll
		for ( int q = 0; q < cfMutableQuad; ++q )
			r[q] = p[q];
	}
	__device__ inline void setQuad( int k, int q, float4 v ) const
	{
		if ( q >= cfMutableQuad )
			mut[2 * k + q - cfMutableQuad] = v;
		else
			rec[(size_t)base * cfQuads + (size_t)k * cfMutableQuad + q] = v;
	}
	__device__ inline int2 bodies( int k ) const
	{
		return sb[k];
	}
	__device__ inline float4& b0( int i ) const
	{
		return P[2 * i];
	}
	__device__ inline float4& b1( int i ) const
	{
		return P[2 * i + 1];
	}
	__device__ inline float2& mass( int i ) const
	{
		return M[i];
	}
};

// the slot bodies live in q0.x and q0.y of the record (int bits)
struct LdsStore
{
	float4* rec; // record of slot k at rec[k * cfQuads]
	float4* P;
	float2* M;
	// the static body's identity state (v = 0, dp = 0, dq = (1, 0), no mass) at this index
	static constexpr bool kIdentityBody = true;
	int identity;
	__device__ inline float& at( int field, int k ) const
	{
		return reinterpret_cast<float*>( rec + k * cfQuads )[field];
	}
	__device__ inline float4 quad( int k, int q ) const
	{
		return rec[k * cfQuads + q];
	}
	__device__ inline void setQuad( int k, int q, float4 v ) const
	{
		rec[k * cfQuads + q] = v;
	}
	static constexpr bool kPrefetch = false;
	__device__ inline void loadQuads( int k, float4 ( &r )[cfMutableQuad] ) const
	{
#pragma unroll
		for ( int q = 0; q < cfMutableQuad; ++q )
			r[q] = quad( k, q );
	}
	__device__ inline int2 bodies( int k ) const
	{
		float4 q0 = rec[k * cfQuads];
		return make_int2( __builtin_bit_cast( int, q0.x ), __builtin_bit_cast( int, q0.y ) );
	}
	__device__ inline float4& b0( int i ) const
	{
		return P[2 * i];
	}
	__device__ inline float4& b1( int i ) const
	{
		return P[2 * i + 1];
	}
	__device__ inline float2& mass( int i ) const
	{
		return M[i];
	}
};

// A run of overflow levels of an HBM-store item staged in LDS (cached_level_run): the distinct
// bodies its slots touch (state, masses, item-local index), each slot's two bodies as cache
// indices (a static body stays negative), and for contacts each slot's record and point count;
// an open-addressing table maps item-local body indices to cache indices while it is built.
constexpr int kRunCacheBodies = 2 * kLevelCacheSlots;
constexpr int kRunCacheHash = 4 * kLevelCacheSlots;
struct RunCache
{
	float4 P[2 * kRunCacheBodies];
	float4 rec[kLevelCacheSlots * cfQuads];
	float2 M[kRunCacheBodies];
	int key[kRunCacheBodies];
	int2 sb[kLevelCacheSlots];
	int pc[kLevelCacheSlots];
	int hashKey[kRunCacheHash];
	int hashVal[kRunCacheHash];
	int count;
};

// The store a cached run's levels use: slot k (of the run, first slot `base`) and cache body i.
struct RunStore
{
	RunCache* c;
	int base;
	static constexpr bool kIdentityBody = false;
	static constexpr bool kPrefetch = false;
	__device__ inline float& at( int field, int k ) const
	{
		return reinterpret_cast<float*>( c->rec + ( k - base ) * cfQuads )[field];
	}
	__device__ inline float4 quad( int k, int q ) const
	{
		return c->rec[( k - base ) * cfQuads + q];
	}
	__device__ inline void setQuad( int k, int q, float4 v ) const
	{
		c->rec[( k - base ) * cfQuads + q] = v;
	}
	__device__ inline void loadQuads( int k, float4 ( &r )[cfMutableQuad] ) const
	{
#pragma unroll
		for ( int q = 0; q < cfMutableQuad; ++q )
			r[q] = quad( k, q );
	}
	__device__ inline int2 bodies( int k ) const
	{
		return c->sb[k - base];
	}
	__device__ inline float4& b0( int i ) const
	{
		return c->P[2 * i];
	}
	__device__ inline float4& b1( int i ) const
	{
		return c->P[2 * i + 1];
	}
	__device__ inline float2& mass( int i ) const
	{
		return c->M[i];
	}
};

// the slot bodies a record's q0 carries (int bits in x and y)
__device__ inline int2 quad_bodies( float4 q0 )
{
	return make_int2( __builtin_bit_cast( int, q0.x ), __builtin_bit_cast( int, q0.y ) );
}

// inverse mass and inertia of local body i (the static body, -1, has none)
template <class S>
__device__ inline float2 body_mass( const S& C, int i )
{
	if constexpr ( S::kIdentityBody )
		return C.mass( i < 0 ? C.identity : i );
	return i < 0 ? make_float2( 0.0f, 0.0f ) : C.mass( i );
}

// The world's two sets of soft constraint coefficients (biasRate, massScale, impulseScale),
// read once per work item (uniform loads) so that the per-lane choice below is a register
// select rather than a load from a lane-dependent address in the middle of every solve.
struct Softness
{
	float contact[3];
	float stat[3];
};

__device__ inline Softness world_softness( const DevWorld& W )
{
	Softness s;
#pragma unroll
	for ( int i = 0; i < 3; ++i )
	{
		s.contact[i] = W.contactSoft[i];
		s.stat[i] = W.staticSoft[i];
	}
	return s;
}

// soft constraint coefficient i of a contact: the static set when either body is static
// (src/contact_solver.c:650-651)
__device__ inline float contact_soft( const Softness& S, int la, int lb, int i )
{
	return ( la < 0 || lb < 0 ) ? S.stat[i] : S.contact[i];
}

#define PF( name, p ) pointField( cf##name, ( p ) )

// What prepare reads of a contact: its manifold (both point slots; the unused one is ignored)
// and material. Loaded apart from the arithmetic so that a lane can fetch its next slot's
// contact while it prepares the current one.
struct PrepInput
{
	DevManifold m;
	float friction, restitution;
};

__device__ inline PrepInput load_prep( const DevContact* __restrict__ contact )
{
	PrepInput in;
	in.m = contact->manifold;
	in.friction = contact->friction;
	in.restitution = contact->restitution;
	return in;
}

// contact_solver.c:597-808 (lanes) and 21-147 (overflow): one constraint slot
template <class S>
__device__ void prepare_slot( const S& C, int slot, const PrepInput& in, int la, int lb, const DevWorld& W )
{
	const DevManifold& m = in.m;
	// the bodies' inverse masses (equal to the ones the narrowphase copied into the contact)
	float2 MA = body_mass( C, la ), MB = body_mass( C, lb );
	float mA = MA.x, iA = MA.y, mB = MB.x, iB = MB.y;
	float vAx = 0.0f, vAy = 0.0f, wA = 0.0f, vBx = 0.0f, vBy = 0.0f, wB = 0.0f;
	if ( la >= 0 )
	{
		float4 s = C.b0( la );
		vAx = s.x;
		vAy = s.y;
		wA = s.z;
	}
	if ( lb >= 0 )
	{
		float4 s = C.b0( lb );
		vBx = s.x;
		vBy = s.y;
		wB = s.z;
	}
	float nx = m.normalX, ny = m.normalY;
	C.setQuad( slot, 0, make_float4( __builtin_bit_cast( float, la ), __builtin_bit_cast( float, lb ), nx, ny ) );
	float tx = ny, ty = -nx;
	float warm = W.warmStartScale;
	float4 anc[2], pm[2], imp[2];
#pragma unroll
	for ( int j = 0; j < 2; ++j )
	{
		if ( j < m.pointCount )
		{
			const DevManifoldPoint& mp = m.points[j];
			float rAx = mp.anchorAX, rAy = mp.anchorAY, rBx = mp.anchorBX, rBy = mp.anchorBY;
			anc[j] = make_float4( rAx, rAy, rBx, rBy );
			float dx = rBx - rAx, dy = rBy - rAy;
			float baseSeparation = mp.separation - ( dx * nx + dy * ny );
			float rnA = rAx * ny - rAy * nx;
			float rnB = rBx * ny - rBy * nx;
			float kNormal = mA + mB + iA * rnA * rnA + iB * rnB * rnB;
			float normalMass = kNormal > 0.0f ? 1.0f / kNormal : 0.0f;
			float rtA = rAx * ty - rAy * tx;
			float rtB = rBx * ty - rBy * tx;
			float kTangent = mA + mB + iA * rtA * rtA + iB * rtB * rtB;
			float tangentMass = kTangent > 0.0f ? 1.0f / kTangent : 0.0f;
			pm[j] = make_float4( baseSeparation, normalMass, tangentMass, 0.0f );
			float vrAx = vAx + ( -wA ) * rAy, vrAy = vAy + wA * rAx;
			float vrBx = vBx + ( -wB ) * rBy, vrBy = vBy + wB * rBx;
			float relativeVelocity = nx * ( vrBx - vrAx ) + ny * ( vrBy - vrAy );
			imp[j] = make_float4( warm * mp.normalImpulse, warm * mp.tangentImpulse, 0.0f, relativeVelocity );
		}
		else
		{
			anc[j] = make_float4( 0.0f, 0.0f, 0.0f, 0.0f );
			pm[j] = make_float4( 0.0f, 0.0f, 0.0f, 0.0f );
			imp[j] = make_float4( 0.0f, 0.0f, 0.0f, 0.0f );
		}
	}
	pm[0].w = in.friction;
	pm[1].w = in.restitution;
	C.setQuad( slot, 1, anc[0] );
	C.setQuad( slot, 2, anc[1] );
	C.setQuad( slot, 3, pm[0] );
	C.setQuad( slot, 4, pm[1] );
	C.setQuad( slot, 5, imp[0] );
	C.setQuad( slot, 6, imp[1] );
}

__device__ inline void lane_apply( Body& bA, Body& bB, float mA, float iA, float mB, float iB, float rAx, float rAy, float rBx,
								   float rBy, float Px, float Py )
{
	bA.vx = bA.vx - mA * Px;
	bA.vy = bA.vy - mA * Py;
	bA.w = bA.w - iA * ( rAx * Py - rAy * Px );
	bB.vx = bB.vx + mB * Px;
	bB.vy = bB.vy + mB * Py;
	bB.w = bB.w + iB * ( rBx * Py - rBy * Px );
}

// contact_solver.c:810-863
template <class S>
__device__ void lane_warm_start_q( const S& C, int slot, int la, int lb, const float4 ( &r )[cfMutableQuad] )
{
	Body bA = load_body( C, la ), bB = load_body( C, lb );
	float2 MA = body_mass( C, la ), MB = body_mass( C, lb );
	float mA = MA.x, mB = MB.x, iA = MA.y, iB = MB.y;
	const float4 q0 = r[0];
	const float4 anc[2] = { r[1], r[2] };
	const float4 imp[2] = { C.quad( slot, 5 ), C.quad( slot, 6 ) };
	float nx = q0.z, ny = q0.w;
	float tx = ny, ty = 0.0f - nx;
#pragma unroll
	for ( int j = 0; j < 2; ++j )
	{
		float ni = imp[j].x, ti = imp[j].y;
		float rAx = anc[j].x, rAy = anc[j].y;
		float rBx = anc[j].z, rBy = anc[j].w;
		float Px = ni * nx + ti * tx;
		float Py = ni * ny + ti * ty;
		bA.w = bA.w - iA * ( rAx * Py - rAy * Px );
		bA.vx = bA.vx - mA * Px;
		bA.vy = bA.vy - mA * Py;
		bB.w = bB.w + iB * ( rBx * Py - rBy * Px );
		bB.vx = bB.vx + mB * Px;
		bB.vy = bB.vy + mB * Py;
	}
	store_velocity( C, la, bA );
	store_velocity( C, lb, bB );
}

// Whether the colour restitution stage (lane_restitution_q) would leave a slot and its bodies bit
// for bit as they are, whatever its bodies' velocities, given that they are plain
// (body_is_plain); from the slot's read-only quads r and its written quads imp (slot_restitution_
// is_identity reads them from a store). Neither point may
// have restitution mass (the test of src/contact_solver.c:1096-1101: relative velocity above
// -threshold, or no normal impulse) and the inputs must be finite and bounded: then the new
// normal impulse equals the old, the applied impulse is a signed zero, and v - m (+-0) returns v
// for every velocity component that is not -0. The bound (1e15) keeps every product and sum of
// the stage finite.
__device__ inline bool restitution_is_identity( const float4 ( &r )[cfMutableQuad], const float4 ( &imp )[2], float threshold )
{
	const float kBound = 1e15f;
	bool ok = fabsf( r[0].z ) <= kBound && fabsf( r[0].w ) <= kBound && fabsf( r[4].w ) <= kBound;
#pragma unroll
	for ( int j = 0; j < 2; ++j )
	{
		const float4 a = r[1 + j];
		bool noMass = ( imp[j].w + threshold ) > 0.0f || imp[j].z == 0.0f;
		ok = ok && noMass && fabsf( imp[j].x ) <= kBound && fabsf( imp[j].w ) <= kBound;
		ok = ok && fabsf( a.x ) <= kBound && fabsf( a.y ) <= kBound && fabsf( a.z ) <= kBound && fabsf( a.w ) <= kBound;
	}
	return ok;
}

template <class S>
__device__ inline bool slot_restitution_is_identity( const S& C, int k, float threshold )
{
	float4 r[cfMutableQuad];
	C.loadQuads( k, r );
	const float4 imp[2] = { C.quad( k, 5 ), C.quad( k, 6 ) };
	return restitution_is_identity( r, imp, threshold );
}

// contact_solver.c:865-1075
template <class S>
__device__ void lane_solve_q( const S& C, int slot, int la, int lb, const float4 ( &r )[cfMutableQuad], float inv_h, float minBiasVel,
							  bool useBias, const Softness& soft )
{
	Body bA = load_body( C, la ), bB = load_body( C, lb );
	float2 MA = body_mass( C, la ), MB = body_mass( C, lb );
	float mA = MA.x, mB = MB.x, iA = MA.y, iB = MB.y;
	const float4 q0 = r[0];
	const float4 anc[2] = { r[1], r[2] };
	const float4 pm[2] = { r[3], r[4] };
	float4 imp[2] = { C.quad( slot, 5 ), C.quad( slot, 6 ) };
	float nx = q0.z, ny = q0.w;
	float biasRate = useBias ? contact_soft( soft, la, lb, 0 ) : 0.0f;
	float massScale = useBias ? contact_soft( soft, la, lb, 1 ) : 1.0f;
	float impulseScale = useBias ? contact_soft( soft, la, lb, 2 ) : 0.0f;
	float dpx = bB.dpx - bA.dpx, dpy = bB.dpy - bA.dpy;
	float nImp[2];
#pragma unroll
	for ( int j = 0; j < 2; ++j )
	{
		float rAx = anc[j].x, rAy = anc[j].y;
		float rBx = anc[j].z, rBy = anc[j].w;
		float rsAx = bA.dqc * rAx - bA.dqs * rAy, rsAy = bA.dqs * rAx + bA.dqc * rAy;
		float rsBx = bB.dqc * rBx - bB.dqs * rBy, rsBy = bB.dqs * rBx + bB.dqc * rBy;
		float dsx = dpx + ( rsBx - rsAx ), dsy = dpy + ( rsBy - rsAy );
		float s = ( nx * dsx + ny * dsy ) + pm[j].x;
		float specBias = s * inv_h;
		float softBias = maxr( biasRate * s, minBiasVel );
		float bias = s > 0.0f ? specBias : softBias;
		float dvx = ( bB.vx - bB.w * rBy ) - ( bA.vx - bA.w * rAy );
		float dvy = ( bB.vy + bB.w * rBx ) - ( bA.vy + bA.w * rAx );
		float vn = dvx * nx + dvy * ny;
		float ni = imp[j].x;
		float negImpulse = pm[j].y * ( massScale * ( vn + bias ) ) + impulseScale * ni;
		float newImpulse = maxr( ni - negImpulse, 0.0f );
		float impulse = newImpulse - ni;
		imp[j].x = newImpulse;
		imp[j].z = maxr( imp[j].z, newImpulse );
		nImp[j] = newImpulse;
		lane_apply( bA, bB, mA, iA, mB, iB, rAx, rAy, rBx, rBy, impulse * nx, impulse * ny );
	}
	float tx = ny, ty = 0.0f - nx;
	float friction = pm[0].w;
#pragma unroll
	for ( int j = 0; j < 2; ++j )
	{
		float rAx = anc[j].x, rAy = anc[j].y;
		float rBx = anc[j].z, rBy = anc[j].w;
		float dvx = ( bB.vx - bB.w * rBy ) - ( bA.vx - bA.w * rAy );
		float dvy = ( bB.vy + bB.w * rBx ) - ( bA.vy + bA.w * rAx );
		float vt = dvx * tx + dvy * ty;
		float negImpulse = pm[j].z * vt;
		float maxFriction = friction * nImp[j];
		float ti = imp[j].y;
		float newImpulse = ti - negImpulse;
		newImpulse = maxr( 0.0f - maxFriction, minr( newImpulse, maxFriction ) );
		float impulse = newImpulse - ti;
		imp[j].y = newImpulse;
		lane_apply( bA, bB, mA, iA, mB, iB, rAx, rAy, rBx, rBy, impulse * tx, impulse * ty );
	}
	C.setQuad( slot, 5, imp[0] );
	C.setQuad( slot, 6, imp[1] );
	store_velocity( C, la, bA );
	store_velocity( C, lb, bB );
}

// contact_solver.c:1077-1178
template <class S>
__device__ void lane_restitution_q( const S& C, int slot, int la, int lb, const float4 ( &r )[cfMutableQuad], float threshold )
{
	Body bA = load_body( C, la ), bB = load_body( C, lb );
	float2 MA = body_mass( C, la ), MB = body_mass( C, lb );
	float mA = MA.x, mB = MB.x, iA = MA.y, iB = MB.y;
	const float4 q0 = r[0];
	const float4 anc[2] = { r[1], r[2] };
	const float4 pm[2] = { r[3], r[4] };
	float4 imp[2] = { C.quad( slot, 5 ), C.quad( slot, 6 ) };
	float nx = q0.z, ny = q0.w;
	float restitution = pm[1].w;
#pragma unroll
	for ( int j = 0; j < 2; ++j )
	{
		float relVel = imp[j].w;
		bool skip = ( relVel + threshold ) > 0.0f || imp[j].z == 0.0f;
		float mass = skip ? 0.0f : pm[j].y;
		float rAx = anc[j].x, rAy = anc[j].y;
		float rBx = anc[j].z, rBy = anc[j].w;
		float dvx = ( bB.vx - bB.w * rBy ) - ( bA.vx - bA.w * rAy );
		float dvy = ( bB.vy + bB.w * rBx ) - ( bA.vy + bA.w * rAx );
		float vn = dvx * nx + dvy * ny;
		float ni = imp[j].x;
		float negImpulse = mass * ( vn + restitution * relVel );
		float newImpulse = maxr( ni - negImpulse, 0.0f );
		float impulse = newImpulse - ni;
		imp[j].x = newImpulse;
		lane_apply( bA, bB, mA, iA, mB, iB, rAx, rAy, rBx, rBy, impulse * nx, impulse * ny );
	}
	C.setQuad( slot, 5, imp[0] );
	C.setQuad( slot, 6, imp[1] );
	store_velocity( C, la, bA );
	store_velocity( C, lb, bB );
}

// A body whose velocity no signed-zero impulse changes: finite, bounded components, none of
// them -0, and bounded inverse mass and inertia.
template <class S>
__device__ inline bool body_is_plain( const S& C, int i )
{
	const float kBound = 1e15f;
	float4 s0 = C.b0( i );
	float2 m = C.mass( i );
	auto plain = [&]( float x ) { return fabsf( x ) <= kBound && __builtin_bit_cast( uint32_t, x ) != 0x80000000u; };
	return plain( s0.x ) && plain( s0.y ) && plain( s0.z ) && fabsf( m.x ) <= kBound && fabsf( m.y ) <= kBound;
}

// the lane stages with the slot's read-only quads loaded from its store
template <class S>
__device__ inline void lane_warm_start( const S& C, int slot, int la, int lb )
{
	float4 r[cfMutableQuad];
	C.loadQuads( slot, r );
	lane_warm_start_q( C, slot, la, lb, r );
}

template <class S>
__device__ inline void lane_solve( const S& C, int slot, int la, int lb, float inv_h, float minBiasVel, bool useBias,
								   const Softness& soft )
{
	float4 r[cfMutableQuad];
	C.loadQuads( slot, r );
	lane_solve_q( C, slot, la, lb, r, inv_h, minBiasVel, useBias, soft );
}

template <class S>
__device__ inline void lane_restitution( const S& C, int slot, int la, int lb, float threshold )
{
	float4 r[cfMutableQuad];
	C.loadQuads( slot, r );
	lane_restitution_q( C, slot, la, lb, r, threshold );
}

// ---- overflow (serial) stages, contact_solver.c:149-471 ----
template <class S>
__device__ void ovf_warm_start( const S& C, int slot, int la, int lb, int pointCount )
{
	Body bA = load_body( C, la ), bB = load_body( C, lb );
	float2 MA = body_mass( C, la ), MB = body_mass( C, lb );
	float mA = MA.x, mB = MB.x, iA = MA.y, iB = MB.y;
	float nx = C.at( cfNormalX, slot ), ny = C.at( cfNormalY, slot );
	float tx = ny, ty = -nx;
	for ( int j = 0; j < pointCount; ++j )
	{
		float ni = C.at( PF( NormalImpulse, j ), slot ), ti = C.at( PF( TangentImpulse, j ), slot );
		float rAx = C.at( PF( AnchorAX, j ), slot ), rAy = C.at( PF( AnchorAY, j ), slot );
		float rBx = C.at( PF( AnchorBX, j ), slot ), rBy = C.at( PF( AnchorBY, j ), slot );
		float Px = ni * nx + ti * tx;
		float Py = ni * ny + ti * ty;
		bA.w -= iA * ( rAx * Py - rAy * Px );
		bA.vx = bA.vx + ( -mA ) * Px;
		bA.vy = bA.vy + ( -mA ) * Py;
		bB.w += iB * ( rBx * Py - rBy * Px );
		bB.vx = bB.vx + mB * Px;
		bB.vy = bB.vy + mB * Py;
	}
	store_velocity( C, la, bA );
	store_velocity( C, lb, bB );
}

template <class S>
__device__ void ovf_solve( const S& C, int slot, int la, int lb, int pointCount, float inv_h, float pushout,
						   bool useBias, const Softness& soft )
{
	Body bA = load_body( C, la ), bB = load_body( C, lb );
	float2 MA = body_mass( C, la ), MB = body_mass( C, lb );
	float mA = MA.x, mB = MB.x, iA = MA.y, iB = MB.y;
	float nx = C.at( cfNormalX, slot ), ny = C.at( cfNormalY, slot );
	float tx = ny, ty = -nx;
	float dpx = bB.dpx - bA.dpx, dpy = bB.dpy - bA.dpy;
	for ( int j = 0; j < pointCount; ++j )
	{
		float rAx = C.at( PF( AnchorAX, j ), slot ), rAy = C.at( PF( AnchorAY, j ), slot );
		float rBx = C.at( PF( AnchorBX, j ), slot ), rBy = C.at( PF( AnchorBY, j ), slot );
		float rBqx = bB.dqc * rBx - bB.dqs * rBy, rBqy = bB.dqs * rBx + bB.dqc * rBy;
		float rAqx = bA.dqc * rAx - bA.dqs * rAy, rAqy = bA.dqs * rAx + bA.dqc * rAy;
		float dsx = dpx + ( rBqx - rAqx ), dsy = dpy + ( rBqy - rAqy );
		float s = ( dsx * nx + dsy * ny ) + C.at( PF( BaseSeparation, j ), slot );
		float velocityBias = 0.0f, massScale = 1.0f, impulseScale = 0.0f;
		if ( s > 0.0f )
			velocityBias = s * inv_h;
		else if ( useBias )
		{
			velocityBias = maxr( contact_soft( soft, la, lb, 0 ) * s, -pushout );
			massScale = contact_soft( soft, la, lb, 1 );
			impulseScale = contact_soft( soft, la, lb, 2 );
		}
		float vrAx = bA.vx + ( -bA.w ) * rAy, vrAy = bA.vy + bA.w * rAx;
		float vrBx = bB.vx + ( -bB.w ) * rBy, vrBy = bB.vy + bB.w * rBx;
		float vn = ( vrBx - vrAx ) * nx + ( vrBy - vrAy ) * ny;
		float ni = C.at( PF( NormalImpulse, j ), slot );
		float impulse = -C.at( PF( NormalMass, j ), slot ) * massScale * ( vn + velocityBias ) - impulseScale * ni;
		float newImpulse = maxr( ni + impulse, 0.0f );
		impulse = newImpulse - ni;
		C.at( PF( NormalImpulse, j ), slot ) = newImpulse;
		C.at( PF( MaxNormalImpulse, j ), slot ) = maxr( C.at( PF( MaxNormalImpulse, j ), slot ), impulse );
		float Px = impulse * nx, Py = impulse * ny;
		bA.vx = bA.vx - mA * Px;
		bA.vy = bA.vy - mA * Py;
		bA.w -= iA * ( rAx * Py - rAy * Px );
		bB.vx = bB.vx + mB * Px;
		bB.vy = bB.vy + mB * Py;
		bB.w += iB * ( rBx * Py - rBy * Px );
	}
	float friction = C.at( cfFriction, slot );
	for ( int j = 0; j < pointCount; ++j )
	{
		float rAx = C.at( PF( AnchorAX, j ), slot ), rAy = C.at( PF( AnchorAY, j ), slot );
		float rBx = C.at( PF( AnchorBX, j ), slot ), rBy = C.at( PF( AnchorBY, j ), slot );
		float vrBx = bB.vx + ( -bB.w ) * rBy, vrBy = bB.vy + bB.w * rBx;
		float vrAx = bA.vx + ( -bA.w ) * rAy, vrAy = bA.vy + bA.w * rAx;
		float vt = ( vrBx - vrAx ) * tx + ( vrBy - vrAy ) * ty;
		float impulse = C.at( PF( TangentMass, j ), slot ) * ( -vt );
		float maxFriction = friction * C.at( PF( NormalImpulse, j ), slot );
		float ti = C.at( PF( TangentImpulse, j ), slot );
		float sum = ti + impulse;
		float newImpulse = sum < -maxFriction ? -maxFriction : ( sum > maxFriction ? maxFriction : sum );
		impulse = newImpulse - ti;
		C.at( PF( TangentImpulse, j ), slot ) = newImpulse;
		float Px = impulse * tx, Py = impulse * ty;
		bA.vx = bA.vx - mA * Px;
		bA.vy = bA.vy - mA * Py;
		bA.w -= iA * ( rAx * Py - rAy * Px );
		bB.vx = bB.vx + mB * Px;
		bB.vy = bB.vy + mB * Py;
		bB.w += iB * ( rBx * Py - rBy * Px );
	}
	store_velocity( C, la, bA );
	store_velocity( C, lb, bB );
}

template <class S>
__device__ void ovf_restitution( const S& C, int slot, int la, int lb, int pointCount, float threshold )
{
	float restitution = C.at( cfRestitution, slot );
	if ( restitution == 0.0f )
		return;
	Body bA = load_body( C, la ), bB = load_body( C, lb );
	float2 MA = body_mass( C, la ), MB = body_mass( C, lb );
	float mA = MA.x, mB = MB.x, iA = MA.y, iB = MB.y;
	float nx = C.at( cfNormalX, slot ), ny = C.at( cfNormalY, slot );
	for ( int j = 0; j < pointCount; ++j )
	{
		float relVel = C.at( PF( RelativeVelocity, j ), slot );
		if ( relVel > -threshold || C.at( PF( MaxNormalImpulse, j ), slot ) == 0.0f )
			continue;
		float rAx = C.at( PF( AnchorAX, j ), slot ), rAy = C.at( PF( AnchorAY, j ), slot );
		float rBx = C.at( PF( AnchorBX, j ), slot ), rBy = C.at( PF( AnchorBY, j ), slot );
		float vrBx = bB.vx + ( -bB.w ) * rBy, vrBy = bB.vy + bB.w * rBx;
		float vrAx = bA.vx + ( -bA.w ) * rAy, vrAy = bA.vy + bA.w * rAx;
		float vn = ( vrBx - vrAx ) * nx + ( vrBy - vrAy ) * ny;
		float ni = C.at( PF( NormalImpulse, j ), slot );
		float impulse = -C.at( PF( NormalMass, j ), slot ) * ( vn + restitution * relVel );
		float newImpulse = maxr( ni + impulse, 0.0f );
		impulse = newImpulse - ni;
		C.at( PF( NormalImpulse, j ), slot ) = newImpulse;
		C.at( PF( MaxNormalImpulse, j ), slot ) = maxr( C.at( PF( MaxNormalImpulse, j ), slot ), impulse );
		float Px = impulse * nx, Py = impulse * ny;
		bA.vx = bA.vx - mA * Px;
		bA.vy = bA.vy - mA * Py;
		bA.w -= iA * ( rAx * Py - rAy * Px );
		bB.vx = bB.vx + mB * Px;
		bB.vy = bB.vy + mB * Py;
		bB.w += iB * ( rBx * Py - rBy * Px );
	}
	store_velocity( C, la, bA );
	store_velocity( C, lb, bB );
}

// The per-body terms of src/solver.c:34-93 that stay the same over a step: the damping factors
// and the velocity deltas from gravity, force and torque (computed as the reference does, so
// computing them once per step gives the same floats as once per substep).
struct VelocityTerms
{
	float linearDamping, angularDamping;
	float dvx, dvy, dw;
};

__device__ inline VelocityTerms velocity_terms( const DevBodySim& sim, const DevWorld& W )
{
	float h = W.h;
	VelocityTerms t;
	t.linearDamping = 1.0f / ( 1.0f + h * sim.linearDamping );
	t.angularDamping = 1.0f / ( 1.0f + h * sim.angularDamping );
	float mg = sim.mass * sim.gravityScale;
	float fx = sim.force[0] + mg * W.gravityX, fy = sim.force[1] + mg * W.gravityY;
	float hm = h * sim.invMass;
	t.dvx = hm * fx;
	t.dvy = hm * fy;
	t.dw = h * sim.invI * sim.torque;
	return t;
}

// solver.c:34-93 for one body of the private set, given its step terms
template <class S>
__device__ void integrate_velocity_t( const S& C, int i, const VelocityTerms& t, const DevWorld& W )
{
	float maxLinearSpeed = W.maxLinearSpeed, maxAngularSpeed = W.maxAngularSpeed;
	float4 s0 = C.b0( i );
	float vx = s0.x, vy = s0.y, w = s0.z;
	vx = t.dvx + t.linearDamping * vx;
	vy = t.dvy + t.linearDamping * vy;
	w = t.dw + t.angularDamping * w;
	if ( vx * vx + vy * vy > maxLinearSpeed * maxLinearSpeed )
	{
		float ratio = maxLinearSpeed / sqrtf( vx * vx + vy * vy );
		vx = ratio * vx;
		vy = ratio * vy;
	}
	if ( w * w > maxAngularSpeed * maxAngularSpeed )
	{
		float ratio = maxAngularSpeed / ( w < 0 ? -w : w );
		w *= ratio;
	}
	C.b0( i ) = make_float4( vx, vy, w, s0.w );
}

template <class S>
__device__ void integrate_velocity( const S& C, int i, const DevBodySim& sim, const DevWorld& W )
{
	integrate_velocity_t( C, i, velocity_terms( sim, W ), W );
}

// solver.c:146-163
template <class S>
__device__ void integrate_position( const S& C, int i, float h )
{
	float4 s0 = C.b0( i );
	float4 s1 = C.b1( i );
	float da = h * s0.z;
	float c2 = s1.z - da * s1.w;
	float s2 = s1.w + da * s1.z;
	float mag = sqrtf( s2 * s2 + c2 * c2 );
	float inv = mag > 0.0f ? 1.0f / mag : 0.0f;
	s1.z = c2 * inv;
	s1.w = s2 * inv;
	s1.x = s1.x + h * s0.x;
	s1.y = s1.y + h * s0.y;
	C.b1( i ) = s1;
}


} // namespace b2mi
