// Narrowphase manifold functions, compiled for both the GPU (k_collide) and the host (the
// exported b2Collide* functions of the public C API), so one source gives both.
//
// Each function restates src/manifold.c / src/distance.c of the reference with the same
// operation order; compiled with -ffp-contract=off and correctly rounded division and
// square root they return the same IEEE floats as the reference. min/max use the
// reference's x86 operand rule (a > b ? a : b).
#pragma once

#include <hip/hip_runtime.h>

#include <float.h>
#include <stdint.h>

#include "b2mi_layout.h"
#include "box2d/math_functions.h"

#define B2MI_HD __host__ __device__

namespace b2mi
{

// ---- constants of the reference (src/core.h:110-151) ----
#define K_LINEAR_SLOP 0.005f
#define K_SPECULATIVE ( 4.0f * K_LINEAR_SLOP )
#define K_AABB_MARGIN 0.1f
#define K_TIME_TO_SLEEP 0.5f // src/core.h:151

B2MI_HD inline float maxr( float a, float b )
{
	return a > b ? a : b;
}
B2MI_HD inline float minr( float a, float b )
{
	return a < b ? a : b;
}

B2MI_HD inline b2Vec2 v2( float x, float y )
{
	b2Vec2 r = { x, y };
	return r;
}

// src/math_functions.c:55-93
B2MI_HD inline b2Vec2 d_normalize( b2Vec2 v )
{
	float length = b2Length( v );
	if ( length < FLT_EPSILON )
		return v2( 0.0f, 0.0f );
	float inv = 1.0f / length;
	return v2( inv * v.x, inv * v.y );
}

B2MI_HD inline b2Vec2 d_length_normalize( float* length, b2Vec2 v )
{
	*length = b2Length( v );
	if ( *length < FLT_EPSILON )
		return v2( 0.0f, 0.0f );
	float inv = 1.0f / *length;
	return v2( inv * v.x, inv * v.y );
}

// ------------------------------------------------------------------------------------
// narrowphase

// a polygon of at most N vertices (N = 8, b2_maxPolygonVertices; N = 4 for the box and
// capsule pairs that dominate the scenes, with half the loop trips and selects)
template <int N>
struct PolyN
{
	b2Vec2 v[N];
	b2Vec2 n[N];
	float radius;
	int count;
};
using Poly = PolyN<8>;

// Element i of a polygon array. Every polygon array here is indexed either by a constant (fully
// unrolled loops over the 8 slots, guarded by i < count) or through pick(), so the arrays live
// in registers on the GPU instead of per-lane scratch memory.
template <int N>
B2MI_HD inline b2Vec2 pick( const b2Vec2 ( &a )[N], int i )
{
	// bitwise select: a compare-and-select chain would be folded back into one indexed load
	uint32_t x = 0u, y = 0u;
#pragma unroll
	for ( int k = 0; k < N; ++k )
	{
		uint32_t mask = 0u - (uint32_t)( i == k );
		x |= __builtin_bit_cast( uint32_t, a[k].x ) & mask;
		y |= __builtin_bit_cast( uint32_t, a[k].y ) & mask;
	}
	return v2( __builtin_bit_cast( float, x ), __builtin_bit_cast( float, y ) );
}

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

B2MI_HD inline void mf_clear( Manifold& m )
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

B2MI_HD inline void mf_finish_point( Manifold& m, b2Transform xfA, b2Transform xfB, b2Vec2 localNormal, b2Vec2 cpA, float sep )
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
template <int N>
B2MI_HD inline void make_capsule( PolyN<N>& s, b2Vec2 p1, b2Vec2 p2, float radius )
{
#pragma unroll
	for ( int i = 2; i < N; ++i )
	{
		s.v[i] = v2( 0.0f, 0.0f );
		s.n[i] = v2( 0.0f, 0.0f );
	}
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
B2MI_HD inline void collide_circles( Manifold& m, b2Vec2 cA, float rA, b2Transform xfA, b2Vec2 cB, float rB, b2Transform xfB )
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
B2MI_HD inline void collide_capsule_circle( Manifold& m, b2Vec2 p1, b2Vec2 p2, float rA, b2Transform xfA, b2Vec2 cB, float rB,
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
template <int N>
B2MI_HD inline void collide_polygon_circle( Manifold& m, const PolyN<N>& A, b2Transform xfA, b2Vec2 cB, float rB, b2Transform xfB )
{
	b2Transform xf = b2InvMulTransforms( xfA, xfB );
	b2Vec2 c = b2TransformPoint( xf, cB );
	float radiusA = A.radius;
	float radius = radiusA + rB;
	int normalIndex = 0;
	float separation = -FLT_MAX;
#pragma unroll
	for ( int i = 0; i < N; ++i )
	{
		if ( i < A.count )
		{
			float s = b2Dot( A.n[i], b2Sub( c, A.v[i] ) );
			if ( s > separation )
			{
				separation = s;
				normalIndex = i;
			}
		}
	}
	if ( separation > radius + K_SPECULATIVE )
		return;
	int i1 = normalIndex;
	int i2 = i1 + 1 < A.count ? i1 + 1 : 0;
	b2Vec2 v1 = pick( A.v, i1 ), v2_ = pick( A.v, i2 );
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
		b2Vec2 normal = pick( A.n, normalIndex );
		b2Vec2 ca = b2MulAdd( c, radiusA - b2Dot( b2Sub( c, v1 ), normal ), normal );
		b2Vec2 cb = b2MulSub( c, rB, normal );
		mf_finish_point( m, xfA, xfB, normal, b2Lerp( ca, cb, 0.5f ), separation - radius );
	}
}

// manifold.c:414-450
template <int N>
B2MI_HD inline float find_max_separation( int* edgeIndex, const PolyN<N>& p1, const PolyN<N>& p2 )
{
	int bestIndex = 0;
	float maxSeparation = -FLT_MAX;
#pragma unroll
	for ( int i = 0; i < N; ++i )
	{
		if ( i < p1.count )
		{
			b2Vec2 n = p1.n[i];
			b2Vec2 v1 = p1.v[i];
			float si = FLT_MAX;
#pragma unroll
			for ( int j = 0; j < N; ++j )
			{
				if ( j < p2.count )
				{
					float sij = b2Dot( n, b2Sub( p2.v[j], v1 ) );
					if ( sij < si )
						si = sij;
				}
			}
			if ( si > maxSeparation )
			{
				maxSeparation = si;
				bestIndex = i;
			}
		}
	}
	*edgeIndex = bestIndex;
	return maxSeparation;
}

// distance.c:32-105
B2MI_HD inline void segment_distance( b2Vec2 p1, b2Vec2 q1, b2Vec2 p2, b2Vec2 q2, float* f1out, float* f2out, float* distSq )
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
template <int N>
B2MI_HD inline void clip_polygons( Manifold& m, const PolyN<N>& polyA, const PolyN<N>& polyB, int edgeA, int edgeB, bool flip )
{
	int i11, i12, i21, i22;
	b2Vec2 normal, v11, v12, v21, v22;
	float r1, r2;
	if ( flip )
	{
		i11 = edgeB;
		i12 = edgeB + 1 < polyB.count ? edgeB + 1 : 0;
		i21 = edgeA;
		i22 = edgeA + 1 < polyA.count ? edgeA + 1 : 0;
		normal = pick( polyB.n, i11 );
		v11 = pick( polyB.v, i11 );
		v12 = pick( polyB.v, i12 );
		v21 = pick( polyA.v, i21 );
		v22 = pick( polyA.v, i22 );
		r1 = polyB.radius;
		r2 = polyA.radius;
	}
	else
	{
		i11 = edgeA;
		i12 = edgeA + 1 < polyA.count ? edgeA + 1 : 0;
		i21 = edgeB;
		i22 = edgeB + 1 < polyB.count ? edgeB + 1 : 0;
		normal = pick( polyA.n, i11 );
		v11 = pick( polyA.v, i11 );
		v12 = pick( polyA.v, i12 );
		v21 = pick( polyB.v, i21 );
		v22 = pick( polyB.v, i22 );
		r1 = polyA.radius;
		r2 = polyB.radius;
	}
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

B2MI_HD inline bool vertex_point( Manifold& m, b2Vec2 vA, b2Vec2 vB, float distSq, float rA, float rB, float radius, int idA,
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
template <int N>
B2MI_HD inline void collide_polygons( Manifold& m, const PolyN<N>& polygonA, b2Transform xfA, const PolyN<N>& polygonB,
									  b2Transform xfB )
{
	b2Vec2 origin = polygonA.v[0];
	b2Transform sfA;
	sfA.p = b2Add( xfA.p, b2RotateVector( xfA.q, origin ) );
	sfA.q = xfA.q;
	b2Transform xf = b2InvMulTransforms( sfA, xfB );

	PolyN<N> localA;
	localA.count = polygonA.count;
	localA.radius = polygonA.radius;
	localA.v[0] = v2( 0.0f, 0.0f );
	localA.n[0] = polygonA.n[0];
#pragma unroll
	for ( int i = 1; i < N; ++i )
	{
		localA.v[i] = i < localA.count ? b2Sub( polygonA.v[i], origin ) : v2( 0.0f, 0.0f );
		localA.n[i] = polygonA.n[i];
	}
	PolyN<N> localB;
	localB.count = polygonB.count;
	localB.radius = polygonB.radius;
#pragma unroll
	for ( int i = 0; i < N; ++i )
	{
		localB.v[i] = i < localB.count ? b2TransformPoint( xf, polygonB.v[i] ) : v2( 0.0f, 0.0f );
		localB.n[i] = i < localB.count ? b2RotateVector( xf.q, polygonB.n[i] ) : v2( 0.0f, 0.0f );
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
		b2Vec2 searchDirection = pick( localA.n, edgeA );
		edgeB = 0;
		float minDot = FLT_MAX;
#pragma unroll
		for ( int i = 0; i < N; ++i )
		{
			float dot = b2Dot( searchDirection, localB.n[i] );
			if ( i < localB.count && dot < minDot )
			{
				minDot = dot;
				edgeB = i;
			}
		}
	}
	else
	{
		flip = true;
		b2Vec2 searchDirection = pick( localB.n, edgeB );
		edgeA = 0;
		float minDot = FLT_MAX;
#pragma unroll
		for ( int i = 0; i < N; ++i )
		{
			float dot = b2Dot( searchDirection, localA.n[i] );
			if ( i < localA.count && dot < minDot )
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
		b2Vec2 v11 = pick( localA.v, i11 ), v12 = pick( localA.v, i12 );
		b2Vec2 v21 = pick( localB.v, i21 ), v22 = pick( localB.v, i22 );
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
#pragma unroll
		for ( int i = 0; i < 2; ++i )
		{
			if ( i >= m.pointCount )
				break;
			m.anchorA[i] = b2RotateVector( xfA.q, b2Add( m.anchorA[i], origin ) );
			m.anchorB[i] = b2Add( m.anchorA[i], b2Sub( xfA.p, xfB.p ) );
			m.point[i] = b2Add( xfA.p, m.anchorA[i] );
		}
	}
}

template <int N>
B2MI_HD inline void load_poly( PolyN<N>& p, const DevShape& s )
{
	p.count = s.count;
	p.radius = s.radius;
#pragma unroll
	for ( int i = 0; i < N; ++i )
	{
		p.v[i] = i < s.count ? v2( s.vert( i )[0], s.vert( i )[1] ) : v2( 0.0f, 0.0f );
		p.n[i] = i < s.count ? v2( s.vert( i )[2], s.vert( i )[3] ) : v2( 0.0f, 0.0f );
	}
}

// polygon, capsule or segment pairs at N vertex slots
template <int N>
B2MI_HD inline void collide_pair( Manifold& m, const DevShape& A, b2Transform xfA, const DevShape& B, b2Transform xfB, b2Vec2 a0,
								  b2Vec2 a1, b2Vec2 b0, b2Vec2 b1 )
{
	PolyN<N> pa, pb;
	if ( A.type == kPolygon )
		load_poly( pa, A );
	else
		make_capsule( pa, a0, a1, A.type == kSegment ? 0.0f : A.radius );
	if ( B.type == kPolygon )
		load_poly( pb, B );
	else
		make_capsule( pb, b0, b1, B.type == kSegment ? 0.0f : B.radius );
	collide_polygons( m, pa, xfA, pb, xfB );
}

// dispatch of src/contact.c:150-183 (shape A is always the primary type of the pair)
B2MI_HD inline void collide_shapes( Manifold& m, const DevShape& A, b2Transform xfA, const DevShape& B, b2Transform xfB )
{
	mf_clear( m );
	int ta = A.type, tb = B.type;
	b2Vec2 a0 = v2( A.vn[0][0], A.vn[0][1] ), a1 = v2( A.vn[1][0], A.vn[1][1] );
	b2Vec2 b0 = v2( B.vn[0][0], B.vn[0][1] ), b1 = v2( B.vn[1][0], B.vn[1][1] );
	if ( tb == kCircle )
	{
		if ( ta == kCircle )
			collide_circles( m, a0, A.radius, xfA, b0, B.radius, xfB );
		else if ( ta == kCapsule )
			collide_capsule_circle( m, a0, a1, A.radius, xfA, b0, B.radius, xfB );
		else if ( ta == kSegment )
			collide_capsule_circle( m, a0, a1, 0.0f, xfA, b0, B.radius, xfB );
		else if ( A.count <= 4 )
		{
			PolyN<4> pa;
			load_poly( pa, A );
			collide_polygon_circle( m, pa, xfA, b0, B.radius, xfB );
		}
		else
		{
			Poly pa;
			load_poly( pa, A );
			collide_polygon_circle( m, pa, xfA, b0, B.radius, xfB );
		}
		return;
	}
	// the same code at 4 or 8 vertex slots: the arithmetic on the used slots is the same
	int na = ta == kPolygon ? A.count : 2, nb = tb == kPolygon ? B.count : 2;
	if ( na <= 4 && nb <= 4 )
		collide_pair<4>( m, A, xfA, B, xfB, a0, a1, b0, b1 );
	else
		collide_pair<8>( m, A, xfA, B, xfB, a0, a1, b0, b1 );
}

} // namespace b2mi
