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

#include "b2mi_layout.h"
#include "box2d/math_functions.h"

namespace b2mi
{

// ---- constants of the reference (src/core.h:110-151) ----
#define K_LINEAR_SLOP 0.005f
#define K_SPECULATIVE ( 4.0f * K_LINEAR_SLOP )
#define K_AABB_MARGIN 0.1f

__device__ inline float maxr( float a, float b )
{
	return a > b ? a : b;
}
__device__ inline float minr( float a, float b )
{
	return a < b ? a : b;
}

__device__ inline b2Vec2 v2( float x, float y )
{
	b2Vec2 r = { x, y };
	return r;
}

// src/math_functions.c:55-93
__device__ inline b2Vec2 d_normalize( b2Vec2 v )
{
	float length = b2Length( v );
	if ( length < FLT_EPSILON )
		return v2( 0.0f, 0.0f );
	float inv = 1.0f / length;
	return v2( inv * v.x, inv * v.y );
}

__device__ inline b2Vec2 d_length_normalize( float* length, b2Vec2 v )
{
	*length = b2Length( v );
	if ( *length < FLT_EPSILON )
		return v2( 0.0f, 0.0f );
	float inv = 1.0f / *length;
	return v2( inv * v.x, inv * v.y );
}

// ------------------------------------------------------------------------------------
// narrowphase

struct Poly
{
	b2Vec2 v[8];
	b2Vec2 n[8];
	float radius;
	int count;
};

struct Manifold
{
	b2Vec2 normal;
	b2Vec2 anchorA[2];
	b2Vec2 anchorB[2];
	b2Vec2 point[2];
	float separation[2];
	uint16_t id[2];
	int pointCount;
};

__device__ inline void mf_clear( Manifold& m )
{
	m.pointCount = 0;
	m.normal = v2( 0.0f, 0.0f );
	for ( int i = 0; i < 2; ++i )
	{
		m.anchorA[i] = v2( 0.0f, 0.0f );
		m.anchorB[i] = v2( 0.0f, 0.0f );
		m.point[i] = v2( 0.0f, 0.0f );
		m.separation[i] = 0.0f;
		m.id[i] = 0;
	}
}

__device__ inline void mf_finish_point( Manifold& m, b2Transform xfA, b2Transform xfB, b2Vec2 localNormal, b2Vec2 cpA, float sep )
{
	m.normal = b2RotateVector( xfA.q, localNormal );
	m.anchorA[0] = b2RotateVector( xfA.q, cpA );
	m.anchorB[0] = b2Add( m.anchorA[0], b2Sub( xfA.p, xfB.p ) );
	m.point[0] = b2Add( xfA.p, m.anchorA[0] );
	m.separation[0] = sep;
	m.id[0] = 0;
	m.pointCount = 1;
}

// manifold.c:14-30
__device__ inline void make_capsule( Poly& s, b2Vec2 p1, b2Vec2 p2, float radius )
{
	s.v[0] = p1;
	s.v[1] = p2;
	b2Vec2 d = b2Sub( p2, p1 );
	float length = b2Length( d );
	b2Vec2 axis = v2( 0.0f, 0.0f );
	if ( length >= FLT_EPSILON )
	{
		float inv = 1.0f / length;
		axis = v2( inv * d.x, inv * d.y );
	}
	b2Vec2 normal = b2RightPerp( axis );
	s.n[0] = normal;
	s.n[1] = b2Neg( normal );
	s.count = 2;
	s.radius = radius;
}

// manifold.c:36-70
__device__ void collide_circles( Manifold& m, b2Vec2 cA, float rA, b2Transform xfA, b2Vec2 cB, float rB, b2Transform xfB )
{
	b2Transform xf = b2InvMulTransforms( xfA, xfB );
	b2Vec2 pointA = cA;
	b2Vec2 pointB = b2TransformPoint( xf, cB );
	float distance;
	b2Vec2 normal = d_length_normalize( &distance, b2Sub( pointB, pointA ) );
	float separation = distance - rA - rB;
	if ( separation > K_SPECULATIVE )
		return;
	b2Vec2 ca = b2MulAdd( pointA, rA, normal );
	b2Vec2 cb = b2MulAdd( pointB, -rB, normal );
	mf_finish_point( m, xfA, xfB, normal, b2Lerp( ca, cb, 0.5f ), separation );
}

// manifold.c:73-135
__device__ void collide_capsule_circle( Manifold& m, b2Vec2 p1, b2Vec2 p2, float rA, b2Transform xfA, b2Vec2 cB, float rB,
										b2Transform xfB )
{
	b2Transform xf = b2InvMulTransforms( xfA, xfB );
	b2Vec2 pB = b2TransformPoint( xf, cB );
	b2Vec2 e = b2Sub( p2, p1 );
	b2Vec2 pA;
	float s1 = b2Dot( b2Sub( pB, p1 ), e );
	float s2 = b2Dot( b2Sub( p2, pB ), e );
	if ( s1 < 0.0f )
		pA = p1;
	else if ( s2 < 0.0f )
		pA = p2;
	else
	{
		float s = s1 / b2Dot( e, e );
		pA = b2MulAdd( p1, s, e );
	}
	float distance;
	b2Vec2 normal = d_length_normalize( &distance, b2Sub( pB, pA ) );
	float separation = distance - rA - rB;
	if ( separation > K_SPECULATIVE )
		return;
	b2Vec2 ca = b2MulAdd( pA, rA, normal );
	b2Vec2 cb = b2MulAdd( pB, -rB, normal );
	mf_finish_point( m, xfA, xfB, normal, b2Lerp( ca, cb, 0.5f ), separation );
}

// manifold.c:137-253
__device__ void collide_polygon_circle( Manifold& m, const Poly& A, b2Transform xfA, b2Vec2 cB, float rB, b2Transform xfB )
{
	b2Transform xf = b2InvMulTransforms( xfA, xfB );
	b2Vec2 c = b2TransformPoint( xf, cB );
	float radiusA = A.radius;
	float radius = radiusA + rB;
	int normalIndex = 0;
	float separation = -FLT_MAX;
	for ( int i = 0; i < A.count; ++i )
	{
		float s = b2Dot( A.n[i], b2Sub( c, A.v[i] ) );
		if ( s > separation )
		{
			separation = s;
			normalIndex = i;
		}
	}
	if ( separation > radius + K_SPECULATIVE )
		return;
	int i1 = normalIndex;
	int i2 = i1 + 1 < A.count ? i1 + 1 : 0;
	b2Vec2 v1 = A.v[i1], v2_ = A.v[i2];
	float u1 = b2Dot( b2Sub( c, v1 ), b2Sub( v2_, v1 ) );
	float u2 = b2Dot( b2Sub( c, v2_ ), b2Sub( v1, v2_ ) );
	if ( ( u1 < 0.0f || u2 < 0.0f ) && separation > FLT_EPSILON )
	{
		b2Vec2 v = u1 < 0.0f ? v1 : v2_;
		b2Vec2 normal = d_normalize( b2Sub( c, v ) );
		separation = b2Dot( b2Sub( c, v ), normal );
		if ( separation > radius + K_SPECULATIVE )
			return;
		b2Vec2 ca = b2MulAdd( v, radiusA, normal );
		b2Vec2 cb = b2MulSub( c, rB, normal );
		mf_finish_point( m, xfA, xfB, normal, b2Lerp( ca, cb, 0.5f ), b2Dot( b2Sub( cb, ca ), normal ) );
	}
	else
	{
		b2Vec2 normal = A.n[normalIndex];
		b2Vec2 ca = b2MulAdd( c, radiusA - b2Dot( b2Sub( c, v1 ), normal ), normal );
		b2Vec2 cb = b2MulSub( c, rB, normal );
		mf_finish_point( m, xfA, xfB, normal, b2Lerp( ca, cb, 0.5f ), separation - radius );
	}
}

// manifold.c:414-450
__device__ inline float find_max_separation( int* edgeIndex, const Poly& p1, const Poly& p2 )
{
	int bestIndex = 0;
	float maxSeparation = -FLT_MAX;
	for ( int i = 0; i < p1.count; ++i )
	{
		b2Vec2 n = p1.n[i];
		b2Vec2 v1 = p1.v[i];
		float si = FLT_MAX;
		for ( int j = 0; j < p2.count; ++j )
		{
			float sij = b2Dot( n, b2Sub( p2.v[j], v1 ) );
			if ( sij < si )
				si = sij;
		}
		if ( si > maxSeparation )
		{
			maxSeparation = si;
			bestIndex = i;
		}
	}
	*edgeIndex = bestIndex;
	return maxSeparation;
}

// distance.c:32-105
__device__ void segment_distance( b2Vec2 p1, b2Vec2 q1, b2Vec2 p2, b2Vec2 q2, float* f1out, float* f2out, float* distSq )
{
	b2Vec2 d1 = b2Sub( q1, p1 );
	b2Vec2 d2 = b2Sub( q2, p2 );
	b2Vec2 r = b2Sub( p1, p2 );
	float dd1 = b2Dot( d1, d1 );
	float dd2 = b2Dot( d2, d2 );
	float rd2 = b2Dot( r, d2 );
	float rd1 = b2Dot( r, d1 );
	const float epsSqr = FLT_EPSILON * FLT_EPSILON;
	float f1, f2;
	if ( dd1 < epsSqr || dd2 < epsSqr )
	{
		if ( dd1 >= epsSqr )
		{
			f1 = b2ClampFloat( -rd1 / dd1, 0.0f, 1.0f );
			f2 = 0.0f;
		}
		else if ( dd2 >= epsSqr )
		{
			f1 = 0.0f;
			f2 = b2ClampFloat( rd2 / dd2, 0.0f, 1.0f );
		}
		else
		{
			f1 = 0.0f;
			f2 = 0.0f;
		}
	}
	else
	{
		float d12 = b2Dot( d1, d2 );
		float denom = dd1 * dd2 - d12 * d12;
		f1 = 0.0f;
		if ( denom != 0.0f )
			f1 = b2ClampFloat( ( d12 * rd2 - rd1 * dd2 ) / denom, 0.0f, 1.0f );
		f2 = ( d12 * f1 + rd2 ) / dd2;
		if ( f2 < 0.0f )
		{
			f2 = 0.0f;
			f1 = b2ClampFloat( -rd1 / dd1, 0.0f, 1.0f );
		}
		else if ( f2 > 1.0f )
		{
			f2 = 1.0f;
			f1 = b2ClampFloat( ( d12 - rd1 ) / dd1, 0.0f, 1.0f );
		}
	}
	b2Vec2 c1 = b2MulAdd( p1, f1, d1 );
	b2Vec2 c2 = b2MulAdd( p2, f2, d2 );
	*f1out = f1;
	*f2out = f2;
	*distSq = b2DistanceSquared( c1, c2 );
}

// manifold.c:276-411
__device__ void clip_polygons( Manifold& m, const Poly& polyA, const Poly& polyB, int edgeA, int edgeB, bool flip )
{
	const Poly* poly1;
	const Poly* poly2;
	int i11, i12, i21, i22;
	if ( flip )
	{
		poly1 = &polyB;
		poly2 = &polyA;
		i11 = edgeB;
		i12 = edgeB + 1 < polyB.count ? edgeB + 1 : 0;
		i21 = edgeA;
		i22 = edgeA + 1 < polyA.count ? edgeA + 1 : 0;
	}
	else
	{
		poly1 = &polyA;
		poly2 = &polyB;
		i11 = edgeA;
		i12 = edgeA + 1 < polyA.count ? edgeA + 1 : 0;
		i21 = edgeB;
		i22 = edgeB + 1 < polyB.count ? edgeB + 1 : 0;
	}
	b2Vec2 normal = poly1->n[i11];
	b2Vec2 v11 = poly1->v[i11], v12 = poly1->v[i12];
	b2Vec2 v21 = poly2->v[i21], v22 = poly2->v[i22];
	b2Vec2 tangent = b2CrossSV( 1.0f, normal );
	float lower1 = 0.0f;
	float upper1 = b2Dot( b2Sub( v12, v11 ), tangent );
	float upper2 = b2Dot( b2Sub( v21, v11 ), tangent );
	float lower2 = b2Dot( b2Sub( v22, v11 ), tangent );
	b2Vec2 vLower;
	if ( lower2 < lower1 && upper2 - lower2 > FLT_EPSILON )
		vLower = b2Lerp( v22, v21, ( lower1 - lower2 ) / ( upper2 - lower2 ) );
	else
		vLower = v22;
	b2Vec2 vUpper;
	if ( upper2 > upper1 && upper2 - lower2 > FLT_EPSILON )
		vUpper = b2Lerp( v22, v21, ( upper1 - lower2 ) / ( upper2 - lower2 ) );
	else
		vUpper = v21;
	float separationLower = b2Dot( b2Sub( vLower, v11 ), normal );
	float separationUpper = b2Dot( b2Sub( vUpper, v11 ), normal );
	float r1 = poly1->radius, r2 = poly2->radius;
	vLower = b2MulAdd( vLower, 0.5f * ( r1 - r2 - separationLower ), normal );
	vUpper = b2MulAdd( vUpper, 0.5f * ( r1 - r2 - separationUpper ), normal );
	float radius = r1 + r2;
	if ( flip == false )
	{
		m.normal = normal;
		m.anchorA[0] = vLower;
		m.separation[0] = separationLower - radius;
		m.id[0] = (uint16_t)( ( i11 & 0xff ) << 8 | ( i22 & 0xff ) );
		m.anchorA[1] = vUpper;
		m.separation[1] = separationUpper - radius;
		m.id[1] = (uint16_t)( ( i12 & 0xff ) << 8 | ( i21 & 0xff ) );
	}
	else
	{
		m.normal = b2Neg( normal );
		m.anchorA[0] = vUpper;
		m.separation[0] = separationUpper - radius;
		m.id[0] = (uint16_t)( ( i21 & 0xff ) << 8 | ( i12 & 0xff ) );
		m.anchorA[1] = vLower;
		m.separation[1] = separationLower - radius;
		m.id[1] = (uint16_t)( ( i22 & 0xff ) << 8 | ( i11 & 0xff ) );
	}
	m.pointCount = 2;
}

__device__ inline bool vertex_point( Manifold& m, b2Vec2 vA, b2Vec2 vB, float distSq, float rA, float rB, float radius, int idA,
									 int idB )
{
	b2Vec2 normal = b2Sub( vB, vA );
	float distance = sqrtf( distSq );
	if ( distance > K_SPECULATIVE + radius )
		return false;
	float invDistance = 1.0f / distance;
	normal.x *= invDistance;
	normal.y *= invDistance;
	b2Vec2 c1 = b2MulAdd( vA, rA, normal );
	b2Vec2 c2 = b2MulAdd( vB, -rB, normal );
	m.normal = normal;
	m.anchorA[0] = b2Lerp( c1, c2, 0.5f );
	m.separation[0] = distance - radius;
	m.id[0] = (uint16_t)( ( idA & 0xff ) << 8 | ( idB & 0xff ) );
	m.pointCount = 1;
	return true;
}

// manifold.c:470-700
__device__ void collide_polygons( Manifold& m, const Poly& polygonA, b2Transform xfA, const Poly& polygonB, b2Transform xfB )
{
	b2Vec2 origin = polygonA.v[0];
	b2Transform sfA;
	sfA.p = b2Add( xfA.p, b2RotateVector( xfA.q, origin ) );
	sfA.q = xfA.q;
	b2Transform xf = b2InvMulTransforms( sfA, xfB );

	Poly localA;
	localA.count = polygonA.count;
	localA.radius = polygonA.radius;
	localA.v[0] = v2( 0.0f, 0.0f );
	localA.n[0] = polygonA.n[0];
	for ( int i = 1; i < localA.count; ++i )
	{
		localA.v[i] = b2Sub( polygonA.v[i], origin );
		localA.n[i] = polygonA.n[i];
	}
	Poly localB;
	localB.count = polygonB.count;
	localB.radius = polygonB.radius;
	for ( int i = 0; i < localB.count; ++i )
	{
		localB.v[i] = b2TransformPoint( xf, polygonB.v[i] );
		localB.n[i] = b2RotateVector( xf.q, polygonB.n[i] );
	}

	int edgeA = 0;
	float separationA = find_max_separation( &edgeA, localA, localB );
	int edgeB = 0;
	float separationB = find_max_separation( &edgeB, localB, localA );
	float radius = localA.radius + localB.radius;
	if ( separationA > K_SPECULATIVE + radius || separationB > K_SPECULATIVE + radius )
		return;

	bool flip;
	if ( separationA >= separationB )
	{
		flip = false;
		b2Vec2 searchDirection = localA.n[edgeA];
		edgeB = 0;
		float minDot = FLT_MAX;
		for ( int i = 0; i < localB.count; ++i )
		{
			float dot = b2Dot( searchDirection, localB.n[i] );
			if ( dot < minDot )
			{
				minDot = dot;
				edgeB = i;
			}
		}
	}
	else
	{
		flip = true;
		b2Vec2 searchDirection = localB.n[edgeB];
		edgeA = 0;
		float minDot = FLT_MAX;
		for ( int i = 0; i < localA.count; ++i )
		{
			float dot = b2Dot( searchDirection, localA.n[i] );
			if ( dot < minDot )
			{
				minDot = dot;
				edgeA = i;
			}
		}
	}

	if ( separationA > 0.1f * K_LINEAR_SLOP || separationB > 0.1f * K_LINEAR_SLOP )
	{
		int i11 = edgeA;
		int i12 = edgeA + 1 < localA.count ? edgeA + 1 : 0;
		int i21 = edgeB;
		int i22 = edgeB + 1 < localB.count ? edgeB + 1 : 0;
		b2Vec2 v11 = localA.v[i11], v12 = localA.v[i12];
		b2Vec2 v21 = localB.v[i21], v22 = localB.v[i22];
		float f1, f2, distSq;
		segment_distance( v11, v12, v21, v22, &f1, &f2, &distSq );
		float rA = localA.radius, rB = localB.radius;
		if ( f1 == 0.0f && f2 == 0.0f )
		{
			if ( !vertex_point( m, v11, v21, distSq, rA, rB, radius, i11, i21 ) )
				return;
		}
		else if ( f1 == 0.0f && f2 == 1.0f )
		{
			if ( !vertex_point( m, v11, v22, distSq, rA, rB, radius, i11, i22 ) )
				return;
		}
		else if ( f1 == 1.0f && f2 == 0.0f )
		{
			if ( !vertex_point( m, v12, v21, distSq, rA, rB, radius, i12, i21 ) )
				return;
		}
		else if ( f1 == 1.0f && f2 == 1.0f )
		{
			if ( !vertex_point( m, v12, v22, distSq, rA, rB, radius, i12, i22 ) )
				return;
		}
		else
		{
			clip_polygons( m, localA, localB, edgeA, edgeB, flip );
		}
	}
	else
	{
		clip_polygons( m, localA, localB, edgeA, edgeB, flip );
	}

	if ( m.pointCount > 0 )
	{
		m.normal = b2RotateVector( xfA.q, m.normal );
		for ( int i = 0; i < m.pointCount; ++i )
		{
			m.anchorA[i] = b2RotateVector( xfA.q, b2Add( m.anchorA[i], origin ) );
			m.anchorB[i] = b2Add( m.anchorA[i], b2Sub( xfA.p, xfB.p ) );
			m.point[i] = b2Add( xfA.p, m.anchorA[i] );
		}
	}
}

__device__ inline void load_poly( Poly& p, const DevShape& s )
{
	p.count = s.count;
	p.radius = s.radius;
	for ( int i = 0; i < s.count; ++i )
	{
		p.v[i] = v2( s.v[i][0], s.v[i][1] );
		p.n[i] = v2( s.n[i][0], s.n[i][1] );
	}
}

// dispatch of src/contact.c:150-183 (shape A is always the primary type of the pair)
__device__ void collide_shapes( Manifold& m, const DevShape& A, b2Transform xfA, const DevShape& B, b2Transform xfB )
{
	mf_clear( m );
	int ta = A.type, tb = B.type;
	b2Vec2 a0 = v2( A.v[0][0], A.v[0][1] ), a1 = v2( A.v[1][0], A.v[1][1] );
	b2Vec2 b0 = v2( B.v[0][0], B.v[0][1] ), b1 = v2( B.v[1][0], B.v[1][1] );
	if ( tb == kCircle )
	{
		if ( ta == kCircle )
			collide_circles( m, a0, A.radius, xfA, b0, B.radius, xfB );
		else if ( ta == kCapsule )
			collide_capsule_circle( m, a0, a1, A.radius, xfA, b0, B.radius, xfB );
		else if ( ta == kSegment )
			collide_capsule_circle( m, a0, a1, 0.0f, xfA, b0, B.radius, xfB );
		else
		{
			Poly pa;
			load_poly( pa, A );
			collide_polygon_circle( m, pa, xfA, b0, B.radius, xfB );
		}
		return;
	}
	Poly pa, pb;
	if ( ta == kPolygon )
		load_poly( pa, A );
	else
		make_capsule( pa, a0, a1, ta == kSegment ? 0.0f : A.radius );
	if ( tb == kPolygon )
		load_poly( pb, B );
	else
		make_capsule( pb, b0, b1, tb == kSegment ? 0.0f : B.radius );
	collide_polygons( m, pa, xfA, pb, xfB );
}

__device__ inline b2Transform sim_transform( const DevBodySim& s )
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
	z.bodySimIndexA = -1;
	z.bodySimIndexB = -1;
	z.friction = ci.friction;
	z.restitution = ci.restitution;
	*c = z;
}

// src/world.c:304-362 + src/contact.c:531-581 for one contact
__global__ void k_collide( const DevWorld* __restrict__ worlds, const int2* __restrict__ list, int count,
						   ChangeRecord* __restrict__ changes, int* __restrict__ changeCount )
{
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if ( i >= count )
		return;
	int2 item = list[i];
	const DevWorld& W = worlds[item.x];
	DevContact* c = W.contacts + item.y;
	const DevShape& A = W.shapes[c->shapeIdA];
	const DevShape& B = W.shapes[c->shapeIdB];
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
		const DevBodySim& sa = staticA ? W.statics[A.body & ~kStaticRef] : W.sims[A.body];
		const DevBodySim& sb = staticB ? W.statics[B.body & ~kStaticRef] : W.sims[B.body];
		c->bodySimIndexA = staticA ? -1 : A.body;
		c->invMassA = sa.invMass;
		c->invIA = sa.invI;
		c->bodySimIndexB = staticB ? -1 : B.body;
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

		DevManifold old = c->manifold;
		DevManifold nm;
		memset( &nm, 0, sizeof( nm ) );
		nm.normalX = m.normal.x;
		nm.normalY = m.normal.y;
		nm.pointCount = m.pointCount;
		for ( int k = 0; k < 2; ++k )
		{
			DevManifoldPoint& p = nm.points[k];
			p.pointX = m.point[k].x;
			p.pointY = m.point[k].y;
			p.anchorAX = m.anchorA[k].x;
			p.anchorAY = m.anchorA[k].y;
			p.anchorBX = m.anchorB[k].x;
			p.anchorBY = m.anchorB[k].y;
			p.separation = m.separation[k];
			p.id = m.id[k];
		}
		for ( int k = 0; k < m.pointCount; ++k )
		{
			DevManifoldPoint& p = nm.points[k];
			b2Vec2 aA = b2Sub( v2( p.anchorAX, p.anchorAY ), offsetA );
			b2Vec2 aB = b2Sub( v2( p.anchorBX, p.anchorBY ), offsetB );
			p.anchorAX = aA.x;
			p.anchorAY = aA.y;
			p.anchorBX = aB.x;
			p.anchorBY = aB.y;
			p.normalImpulse = 0.0f;
			p.tangentImpulse = 0.0f;
			p.maxNormalImpulse = 0.0f;
			p.normalVelocity = 0.0f;
			p.persisted = 0;
			for ( int j = 0; j < old.pointCount; ++j )
			{
				if ( old.points[j].id == p.id )
				{
					p.normalImpulse = old.points[j].normalImpulse;
					p.tangentImpulse = old.points[j].tangentImpulse;
					p.persisted = 1;
					break;
				}
			}
		}
		c->manifold = nm;
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
		r.world = item.x;
		r.contactId = item.y;
		r.flags = change;
		r.pad = 0;
		changes[slot] = r;
	}
}

// ------------------------------------------------------------------------------------
// solver

struct Body
{
	float vx, vy, w, flags;
	float dpx, dpy, dqc, dqs;
};

__device__ inline Body load_body( const float4* P, int i )
{
	Body b;
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
	float4 s0 = P[2 * i], s1 = P[2 * i + 1];
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

__device__ inline void store_velocity( float4* P, int i, const Body& b )
{
	if ( i < 0 )
		return;
	P[2 * i] = make_float4( b.vx, b.vy, b.w, b.flags );
}

struct Cons
{
	float* f;
	int cap;
	__device__ inline float& at( int field, int slot ) const
	{
		return f[(size_t)field * cap + slot];
	}
};

#define PF( name, p ) ( cf##name + cfPointFields * ( p ) )

// contact_solver.c:597-808 (lanes) and 21-147 (overflow): one constraint slot
__device__ void prepare_slot( const Cons& C, int slot, const DevContact* __restrict__ contact, int la, int lb, const float4* P,
							  const DevWorld& W )
{
	const DevManifold& m = contact->manifold;
	float mA = contact->invMassA, iA = contact->invIA, mB = contact->invMassB, iB = contact->invIB;
	float vAx = 0.0f, vAy = 0.0f, wA = 0.0f, vBx = 0.0f, vBy = 0.0f, wB = 0.0f;
	if ( la >= 0 )
	{
		float4 s = P[2 * la];
		vAx = s.x;
		vAy = s.y;
		wA = s.z;
	}
	if ( lb >= 0 )
	{
		float4 s = P[2 * lb];
		vBx = s.x;
		vBy = s.y;
		wB = s.z;
	}
	C.at( cfInvMassA, slot ) = mA;
	C.at( cfInvMassB, slot ) = mB;
	C.at( cfInvIA, slot ) = iA;
	C.at( cfInvIB, slot ) = iB;
	const float* soft = ( la < 0 || lb < 0 ) ? W.staticSoft : W.contactSoft;
	C.at( cfBiasRate, slot ) = soft[0];
	C.at( cfMassScale, slot ) = soft[1];
	C.at( cfImpulseScale, slot ) = soft[2];
	float nx = m.normalX, ny = m.normalY;
	C.at( cfNormalX, slot ) = nx;
	C.at( cfNormalY, slot ) = ny;
	C.at( cfFriction, slot ) = contact->friction;
	C.at( cfRestitution, slot ) = contact->restitution;
	float tx = ny, ty = -nx;
	float warm = W.warmStartScale;
	for ( int j = 0; j < 2; ++j )
	{
		if ( j < m.pointCount )
		{
			const DevManifoldPoint& mp = m.points[j];
			float rAx = mp.anchorAX, rAy = mp.anchorAY, rBx = mp.anchorBX, rBy = mp.anchorBY;
			C.at( PF( AnchorAX, j ), slot ) = rAx;
			C.at( PF( AnchorAY, j ), slot ) = rAy;
			C.at( PF( AnchorBX, j ), slot ) = rBx;
			C.at( PF( AnchorBY, j ), slot ) = rBy;
			float dx = rBx - rAx, dy = rBy - rAy;
			C.at( PF( BaseSeparation, j ), slot ) = mp.separation - ( dx * nx + dy * ny );
			C.at( PF( NormalImpulse, j ), slot ) = warm * mp.normalImpulse;
			C.at( PF( TangentImpulse, j ), slot ) = warm * mp.tangentImpulse;
			C.at( PF( MaxNormalImpulse, j ), slot ) = 0.0f;
			float rnA = rAx * ny - rAy * nx;
			float rnB = rBx * ny - rBy * nx;
			float kNormal = mA + mB + iA * rnA * rnA + iB * rnB * rnB;
			C.at( PF( NormalMass, j ), slot ) = kNormal > 0.0f ? 1.0f / kNormal : 0.0f;
			float rtA = rAx * ty - rAy * tx;
			float rtB = rBx * ty - rBy * tx;
			float kTangent = mA + mB + iA * rtA * rtA + iB * rtB * rtB;
			C.at( PF( TangentMass, j ), slot ) = kTangent > 0.0f ? 1.0f / kTangent : 0.0f;
			float vrAx = vAx + ( -wA ) * rAy, vrAy = vAy + wA * rAx;
			float vrBx = vBx + ( -wB ) * rBy, vrBy = vBy + wB * rBx;
			C.at( PF( RelativeVelocity, j ), slot ) = nx * ( vrBx - vrAx ) + ny * ( vrBy - vrAy );
		}
		else
		{
			for ( int f = 0; f < cfPointFields; ++f )
				C.at( cfAnchorAX + cfPointFields * j + f, slot ) = 0.0f;
		}
	}
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
__device__ void lane_warm_start( const Cons& C, int slot, int la, int lb, float4* P )
{
	Body bA = load_body( P, la ), bB = load_body( P, lb );
	float mA = C.at( cfInvMassA, slot ), mB = C.at( cfInvMassB, slot ), iA = C.at( cfInvIA, slot ), iB = C.at( cfInvIB, slot );
	float nx = C.at( cfNormalX, slot ), ny = C.at( cfNormalY, slot );
	float tx = ny, ty = 0.0f - nx;
	for ( int j = 0; j < 2; ++j )
	{
		float ni = C.at( PF( NormalImpulse, j ), slot ), ti = C.at( PF( TangentImpulse, j ), slot );
		float rAx = C.at( PF( AnchorAX, j ), slot ), rAy = C.at( PF( AnchorAY, j ), slot );
		float rBx = C.at( PF( AnchorBX, j ), slot ), rBy = C.at( PF( AnchorBY, j ), slot );
		float Px = ni * nx + ti * tx;
		float Py = ni * ny + ti * ty;
		bA.w = bA.w - iA * ( rAx * Py - rAy * Px );
		bA.vx = bA.vx - mA * Px;
		bA.vy = bA.vy - mA * Py;
		bB.w = bB.w + iB * ( rBx * Py - rBy * Px );
		bB.vx = bB.vx + mB * Px;
		bB.vy = bB.vy + mB * Py;
	}
	store_velocity( P, la, bA );
	store_velocity( P, lb, bB );
}

// contact_solver.c:865-1075
__device__ void lane_solve( const Cons& C, int slot, int la, int lb, float4* P, float inv_h, float minBiasVel, bool useBias )
{
	Body bA = load_body( P, la ), bB = load_body( P, lb );
	float mA = C.at( cfInvMassA, slot ), mB = C.at( cfInvMassB, slot ), iA = C.at( cfInvIA, slot ), iB = C.at( cfInvIB, slot );
	float nx = C.at( cfNormalX, slot ), ny = C.at( cfNormalY, slot );
	float biasRate = useBias ? C.at( cfBiasRate, slot ) : 0.0f;
	float massScale = useBias ? C.at( cfMassScale, slot ) : 1.0f;
	float impulseScale = useBias ? C.at( cfImpulseScale, slot ) : 0.0f;
	float dpx = bB.dpx - bA.dpx, dpy = bB.dpy - bA.dpy;
	float nImp[2];
	for ( int j = 0; j < 2; ++j )
	{
		float rAx = C.at( PF( AnchorAX, j ), slot ), rAy = C.at( PF( AnchorAY, j ), slot );
		float rBx = C.at( PF( AnchorBX, j ), slot ), rBy = C.at( PF( AnchorBY, j ), slot );
		float rsAx = bA.dqc * rAx - bA.dqs * rAy, rsAy = bA.dqs * rAx + bA.dqc * rAy;
		float rsBx = bB.dqc * rBx - bB.dqs * rBy, rsBy = bB.dqs * rBx + bB.dqc * rBy;
		float dsx = dpx + ( rsBx - rsAx ), dsy = dpy + ( rsBy - rsAy );
		float s = ( nx * dsx + ny * dsy ) + C.at( PF( BaseSeparation, j ), slot );
		float specBias = s * inv_h;
		float softBias = maxr( biasRate * s, minBiasVel );
		float bias = s > 0.0f ? specBias : softBias;
		float dvx = ( bB.vx - bB.w * rBy ) - ( bA.vx - bA.w * rAy );
		float dvy = ( bB.vy + bB.w * rBx ) - ( bA.vy + bA.w * rAx );
		float vn = dvx * nx + dvy * ny;
		float ni = C.at( PF( NormalImpulse, j ), slot );
		float negImpulse = C.at( PF( NormalMass, j ), slot ) * ( massScale * ( vn + bias ) ) + impulseScale * ni;
		float newImpulse = maxr( ni - negImpulse, 0.0f );
		float impulse = newImpulse - ni;
		C.at( PF( NormalImpulse, j ), slot ) = newImpulse;
		C.at( PF( MaxNormalImpulse, j ), slot ) = maxr( C.at( PF( MaxNormalImpulse, j ), slot ), newImpulse );
		nImp[j] = newImpulse;
		lane_apply( bA, bB, mA, iA, mB, iB, rAx, rAy, rBx, rBy, impulse * nx, impulse * ny );
	}
	float tx = ny, ty = 0.0f - nx;
	float friction = C.at( cfFriction, slot );
	for ( int j = 0; j < 2; ++j )
	{
		float rAx = C.at( PF( AnchorAX, j ), slot ), rAy = C.at( PF( AnchorAY, j ), slot );
		float rBx = C.at( PF( AnchorBX, j ), slot ), rBy = C.at( PF( AnchorBY, j ), slot );
		float dvx = ( bB.vx - bB.w * rBy ) - ( bA.vx - bA.w * rAy );
		float dvy = ( bB.vy + bB.w * rBx ) - ( bA.vy + bA.w * rAx );
		float vt = dvx * tx + dvy * ty;
		float negImpulse = C.at( PF( TangentMass, j ), slot ) * vt;
		float maxFriction = friction * nImp[j];
		float ti = C.at( PF( TangentImpulse, j ), slot );
		float newImpulse = ti - negImpulse;
		newImpulse = maxr( 0.0f - maxFriction, minr( newImpulse, maxFriction ) );
		float impulse = newImpulse - ti;
		C.at( PF( TangentImpulse, j ), slot ) = newImpulse;
		lane_apply( bA, bB, mA, iA, mB, iB, rAx, rAy, rBx, rBy, impulse * tx, impulse * ty );
	}
	store_velocity( P, la, bA );
	store_velocity( P, lb, bB );
}

// contact_solver.c:1077-1178
__device__ void lane_restitution( const Cons& C, int slot, int la, int lb, float4* P, float threshold )
{
	Body bA = load_body( P, la ), bB = load_body( P, lb );
	float mA = C.at( cfInvMassA, slot ), mB = C.at( cfInvMassB, slot ), iA = C.at( cfInvIA, slot ), iB = C.at( cfInvIB, slot );
	float nx = C.at( cfNormalX, slot ), ny = C.at( cfNormalY, slot );
	float restitution = C.at( cfRestitution, slot );
	for ( int j = 0; j < 2; ++j )
	{
		float relVel = C.at( PF( RelativeVelocity, j ), slot );
		bool skip = ( relVel + threshold ) > 0.0f || C.at( PF( MaxNormalImpulse, j ), slot ) == 0.0f;
		float mass = skip ? 0.0f : C.at( PF( NormalMass, j ), slot );
		float rAx = C.at( PF( AnchorAX, j ), slot ), rAy = C.at( PF( AnchorAY, j ), slot );
		float rBx = C.at( PF( AnchorBX, j ), slot ), rBy = C.at( PF( AnchorBY, j ), slot );
		float dvx = ( bB.vx - bB.w * rBy ) - ( bA.vx - bA.w * rAy );
		float dvy = ( bB.vy + bB.w * rBx ) - ( bA.vy + bA.w * rAx );
		float vn = dvx * nx + dvy * ny;
		float ni = C.at( PF( NormalImpulse, j ), slot );
		float negImpulse = mass * ( vn + restitution * relVel );
		float newImpulse = maxr( ni - negImpulse, 0.0f );
		float impulse = newImpulse - ni;
		C.at( PF( NormalImpulse, j ), slot ) = newImpulse;
		lane_apply( bA, bB, mA, iA, mB, iB, rAx, rAy, rBx, rBy, impulse * nx, impulse * ny );
	}
	store_velocity( P, la, bA );
	store_velocity( P, lb, bB );
}

// ---- overflow (serial) stages, contact_solver.c:149-471 ----
__device__ void ovf_warm_start( const Cons& C, int slot, int la, int lb, float4* P, int pointCount )
{
	Body bA = load_body( P, la ), bB = load_body( P, lb );
	float mA = C.at( cfInvMassA, slot ), mB = C.at( cfInvMassB, slot ), iA = C.at( cfInvIA, slot ), iB = C.at( cfInvIB, slot );
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
	store_velocity( P, la, bA );
	store_velocity( P, lb, bB );
}

__device__ void ovf_solve( const Cons& C, int slot, int la, int lb, float4* P, int pointCount, float inv_h, float pushout,
						   bool useBias )
{
	Body bA = load_body( P, la ), bB = load_body( P, lb );
	float mA = C.at( cfInvMassA, slot ), mB = C.at( cfInvMassB, slot ), iA = C.at( cfInvIA, slot ), iB = C.at( cfInvIB, slot );
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
			velocityBias = maxr( C.at( cfBiasRate, slot ) * s, -pushout );
			massScale = C.at( cfMassScale, slot );
			impulseScale = C.at( cfImpulseScale, slot );
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
	store_velocity( P, la, bA );
	store_velocity( P, lb, bB );
}

__device__ void ovf_restitution( const Cons& C, int slot, int la, int lb, float4* P, int pointCount, float threshold )
{
	float restitution = C.at( cfRestitution, slot );
	if ( restitution == 0.0f )
		return;
	Body bA = load_body( P, la ), bB = load_body( P, lb );
	float mA = C.at( cfInvMassA, slot ), mB = C.at( cfInvMassB, slot ), iA = C.at( cfInvIA, slot ), iB = C.at( cfInvIB, slot );
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
	store_velocity( P, la, bA );
	store_velocity( P, lb, bB );
}

// solver.c:34-93 for one body of the private set
__device__ void integrate_velocity( float4* P, int i, const DevBodySim& sim, const DevWorld& W )
{
	float h = W.h;
	float maxLinearSpeed = W.maxLinearSpeed, maxAngularSpeed = W.maxAngularSpeed;
	float4 s0 = P[2 * i];
	float vx = s0.x, vy = s0.y, w = s0.z;
	float linearDamping = 1.0f / ( 1.0f + h * sim.linearDamping );
	float angularDamping = 1.0f / ( 1.0f + h * sim.angularDamping );
	float mg = sim.mass * sim.gravityScale;
	float fx = sim.force[0] + mg * W.gravityX, fy = sim.force[1] + mg * W.gravityY;
	float hm = h * sim.invMass;
	float lvdx = hm * fx, lvdy = hm * fy;
	float avd = h * sim.invI * sim.torque;
	vx = lvdx + linearDamping * vx;
	vy = lvdy + linearDamping * vy;
	w = avd + angularDamping * w;
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
	P[2 * i] = make_float4( vx, vy, w, s0.w );
}

// solver.c:146-163
__device__ void integrate_position( float4* P, int i, float h )
{
	float4 s0 = P[2 * i];
	float4 s1 = P[2 * i + 1];
	float da = h * s0.z;
	float c2 = s1.z - da * s1.w;
	float s2 = s1.w + da * s1.z;
	float mag = sqrtf( s2 * s2 + c2 * c2 );
	float inv = mag > 0.0f ? 1.0f / mag : 0.0f;
	s1.z = c2 * inv;
	s1.w = s2 * inv;
	s1.x = s1.x + h * s0.x;
	s1.y = s1.y + h * s0.y;
	P[2 * i + 1] = s1;
}

__device__ inline void shape_aabb( const DevShape& s, b2Transform xf, float* out )
{
	// src/shape.c:457-478 / src/geometry.c:376-429
	b2Vec2 lower, upper;
	if ( s.type == kCircle )
	{
		b2Vec2 p = b2TransformPoint( xf, v2( s.v[0][0], s.v[0][1] ) );
		float r = s.radius;
		lower = v2( p.x - r, p.y - r );
		upper = v2( p.x + r, p.y + r );
	}
	else if ( s.type == kCapsule || s.type == kSegment )
	{
		b2Vec2 a = b2TransformPoint( xf, v2( s.v[0][0], s.v[0][1] ) );
		b2Vec2 b = b2TransformPoint( xf, v2( s.v[1][0], s.v[1][1] ) );
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
		lower = b2TransformPoint( xf, v2( s.v[0][0], s.v[0][1] ) );
		upper = lower;
		for ( int i = 1; i < s.count; ++i )
		{
			b2Vec2 v = b2TransformPoint( xf, v2( s.v[i][0], s.v[i][1] ) );
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

// solver.c:165-351 (+ the no-hit branch of continuous, solver.c:1035-1066) for one body
__device__ void finalize_body( const DevWorld& W, int simIndex, float4* P, int li, int worldIndex, EnlargeRecord* enl,
							   int* enlCount, int enlCapacity )
{
	DevBodySim& sim = W.sims[simIndex];
	float4 s0 = P[2 * li], s1 = P[2 * li + 1];
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
	sim.force[0] = 0.0f;
	sim.force[1] = 0.0f;
	sim.torque = 0.0f;

	bool isFast = false;
	bool enableSleep = ( W.flags & kWorldSleep ) != 0 && ( sim.flags & kBodyEnableSleep ) != 0;
	if ( enableSleep == false || sleepVelocity > sim.sleepThreshold )
	{
		const float safetyFactor = 0.5f;
		if ( ( sim.flags & kBodyDynamic ) && ( W.flags & kWorldContinuous ) && maxVelocity * W.dt > safetyFactor * sim.minExtent )
			isFast = true;
	}
	// center0/rotation0: immediately for normal bodies, after the (no-hit) sweep check for fast ones
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
		uint32_t flags = 0;
		float box[4] = { 0.0f, 0.0f, 0.0f, 0.0f };
		if ( isFast )
		{
			float box2[4];
			shape_aabb( shape, xf, box2 );
			box[0] = minr( shape.aabb[0], box2[0] );
			box[1] = minr( shape.aabb[1], box2[1] );
			box[2] = maxr( shape.aabb[2], box2[2] );
			box[3] = maxr( shape.aabb[3], box2[3] );
			for ( int c = 0; c < 4; ++c )
				shape.aabb[c] = box2[c];
			flags = kFast | ( ( sim.flags & kBodyBullet ) ? kBulletFast : 0u );
			if ( aabb_contains( shape.fatAABB, box2 ) == false )
			{
				shape.fatAABB[0] = box2[0] - K_AABB_MARGIN;
				shape.fatAABB[1] = box2[1] - K_AABB_MARGIN;
				shape.fatAABB[2] = box2[2] + K_AABB_MARGIN;
				shape.fatAABB[3] = box2[3] + K_AABB_MARGIN;
				flags |= kEnlarged;
			}
		}
		else
		{
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
				flags |= kEnlarged;
			}
		}
		if ( flags )
		{
			int slot = atomicAdd( enlCount, 1 );
			if ( slot < enlCapacity )
			{
				EnlargeRecord r;
				r.world = worldIndex;
				r.simIndex = simIndex;
				r.shapeId = shapeId;
				r.flags = flags;
				for ( int c = 0; c < 4; ++c )
				{
					r.fatAABB[c] = shape.fatAABB[c];
					r.sweepBox[c] = box[c];
				}
				enl[slot] = r;
			}
		}
	}
	// reset deltas, keep velocity
	P[2 * li + 1] = make_float4( 0.0f, 0.0f, 1.0f, 0.0f );
}

// One workgroup per work item; all stages of the step with workgroup barriers.
__global__ void __launch_bounds__( 1024 ) k_solve( const DevWorld* __restrict__ worlds, const WorkItem* __restrict__ items,
												   const int* __restrict__ bodyList, const int* __restrict__ slotContact,
												   const int2* __restrict__ slotBodies, float* cfBase, int cfCapacity,
												   float4* priv, EnlargeRecord* enl, int* enlCount, int enlCapacity )
{
	const WorkItem wi = items[blockIdx.x];
	const DevWorld& W = worlds[wi.world];
	const int tid = threadIdx.x, nt = blockDim.x;
	float4* P = priv + 2 * (size_t)wi.bodyStart;
	const int* bl = bodyList + wi.bodyStart;
	const float4* gstate = reinterpret_cast<const float4*>( W.state );
	Cons C;
	C.f = cfBase;
	C.cap = cfCapacity;
	const int s0 = wi.slotStart;
	const int total = wi.groupStart[kGroups];

	// stage private body state
	for ( int i = tid; i < wi.bodyCount; i += nt )
	{
		int b = bl[i];
		P[2 * i] = gstate[2 * b];
		P[2 * i + 1] = gstate[2 * b + 1];
	}
	__syncthreads();

	// prepare (all groups; the overflow group uses the same per-point formulas)
	for ( int k = tid; k < total; k += nt )
	{
		int slot = s0 + k;
		int2 lb = slotBodies[slot];
		prepare_slot( C, slot, W.contacts + slotContact[slot], lb.x, lb.y, P, W );
	}
	__syncthreads();

	const int ovfBegin = s0 + wi.groupStart[0], ovfEnd = s0 + wi.groupStart[1];
	const bool hasOverflow = ovfEnd > ovfBegin;
	const float inv_h = W.inv_h, h = W.h;
	const float pushout = W.pushout, minBiasVel = -W.pushout;
	const bool warm = ( W.flags & kWorldWarmStart ) != 0;

	for ( int sub = 0; sub < W.subSteps; ++sub )
	{
		for ( int i = tid; i < wi.bodyCount; i += nt )
			integrate_velocity( P, i, W.sims[bl[i]], W );
		__syncthreads();

		// warm start: overflow (always, src/solver.c:639), then colours if enabled
		if ( hasOverflow )
		{
			if ( tid == 0 )
				for ( int slot = ovfBegin; slot < ovfEnd; ++slot )
				{
					int2 lb = slotBodies[slot];
					ovf_warm_start( C, slot, lb.x, lb.y, P, W.contacts[slotContact[slot]].manifold.pointCount );
				}
			__syncthreads();
		}
		if ( warm )
			for ( int g = 1; g < kGroups; ++g )
			{
				int b = s0 + wi.groupStart[g], e = s0 + wi.groupStart[g + 1];
				if ( b == e )
					continue;
				for ( int slot = b + tid; slot < e; slot += nt )
				{
					int2 lb = slotBodies[slot];
					lane_warm_start( C, slot, lb.x, lb.y, P );
				}
				__syncthreads();
			}

		// solve with bias
		if ( hasOverflow )
		{
			if ( tid == 0 )
				for ( int slot = ovfBegin; slot < ovfEnd; ++slot )
				{
					int2 lb = slotBodies[slot];
					ovf_solve( C, slot, lb.x, lb.y, P, W.contacts[slotContact[slot]].manifold.pointCount, inv_h, pushout, true );
				}
			__syncthreads();
		}
		for ( int g = 1; g < kGroups; ++g )
		{
			int b = s0 + wi.groupStart[g], e = s0 + wi.groupStart[g + 1];
			if ( b == e )
				continue;
			for ( int slot = b + tid; slot < e; slot += nt )
			{
				int2 lb = slotBodies[slot];
				lane_solve( C, slot, lb.x, lb.y, P, inv_h, minBiasVel, true );
			}
			__syncthreads();
		}

		for ( int i = tid; i < wi.bodyCount; i += nt )
			integrate_position( P, i, h );
		__syncthreads();

		// relax
		if ( hasOverflow )
		{
			if ( tid == 0 )
				for ( int slot = ovfBegin; slot < ovfEnd; ++slot )
				{
					int2 lb = slotBodies[slot];
					ovf_solve( C, slot, lb.x, lb.y, P, W.contacts[slotContact[slot]].manifold.pointCount, inv_h, pushout, false );
				}
			__syncthreads();
		}
		for ( int g = 1; g < kGroups; ++g )
		{
			int b = s0 + wi.groupStart[g], e = s0 + wi.groupStart[g + 1];
			if ( b == e )
				continue;
			for ( int slot = b + tid; slot < e; slot += nt )
			{
				int2 lb = slotBodies[slot];
				lane_solve( C, slot, lb.x, lb.y, P, inv_h, minBiasVel, false );
			}
			__syncthreads();
		}
	}

	// restitution
	const float threshold = W.restitutionThreshold;
	if ( hasOverflow )
	{
		if ( tid == 0 )
			for ( int slot = ovfBegin; slot < ovfEnd; ++slot )
			{
				int2 lb = slotBodies[slot];
				ovf_restitution( C, slot, lb.x, lb.y, P, W.contacts[slotContact[slot]].manifold.pointCount, threshold );
			}
		__syncthreads();
	}
	for ( int g = 1; g < kGroups; ++g )
	{
		int b = s0 + wi.groupStart[g], e = s0 + wi.groupStart[g + 1];
		if ( b == e )
			continue;
		for ( int slot = b + tid; slot < e; slot += nt )
		{
			int2 lb = slotBodies[slot];
			lane_restitution( C, slot, lb.x, lb.y, P, threshold );
		}
		__syncthreads();
	}

	// store impulses (contact_solver.c:442-471 overflow: pointCount points; 1180-1293 lanes: both)
	for ( int k = tid; k < total; k += nt )
	{
		int slot = s0 + k;
		DevManifold& m = W.contacts[slotContact[slot]].manifold;
		int n = slot < ovfEnd ? m.pointCount : 2;
		for ( int j = 0; j < n; ++j )
		{
			m.points[j].normalImpulse = C.at( PF( NormalImpulse, j ), slot );
			m.points[j].tangentImpulse = C.at( PF( TangentImpulse, j ), slot );
			m.points[j].maxNormalImpulse = C.at( PF( MaxNormalImpulse, j ), slot );
			m.points[j].normalVelocity = C.at( PF( RelativeVelocity, j ), slot );
		}
	}

	// finalize owned bodies and write their state back
	float4* gs = reinterpret_cast<float4*>( W.state );
	for ( int i = tid; i < wi.ownedCount; i += nt )
	{
		int b = bl[i];
		finalize_body( W, b, P, i, wi.world, enl, enlCount, enlCapacity );
		gs[2 * b] = P[2 * i];
		gs[2 * b + 1] = P[2 * i + 1];
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

hipError_t launch_collide( const DevWorld* worlds, const int2* list, int count, ChangeRecord* changes, int* changeCount,
						   hipStream_t stream )
{
	if ( count <= 0 )
		return hipSuccess;
	int block = 256, grid = ( count + block - 1 ) / block;
	hipLaunchKernelGGL( k_collide, dim3( grid ), dim3( block ), 0, stream, worlds, list, count, changes, changeCount );
	return hipGetLastError();
}

hipError_t launch_solve( const DevWorld* worlds, const WorkItem* items, int itemCount, int blockSize, const int* bodyList,
						 const int* slotContact, const int2* slotBodies, float* cf, int cfCapacity, float4* priv,
						 EnlargeRecord* enl, int* enlCount, int enlCapacity, hipStream_t stream )
{
	if ( itemCount <= 0 )
		return hipSuccess;
	hipLaunchKernelGGL( k_solve, dim3( itemCount ), dim3( blockSize ), 0, stream, worlds, items, bodyList, slotContact, slotBodies,
						cf, cfCapacity, priv, enl, enlCount, enlCapacity );
	return hipGetLastError();
}

} // namespace b2mi
