// Time of impact for the continuous stage: GJK distance between two point clouds and the local
// separating-axis root finder, shared by the continuous kernel (kernels.hip) and the host exports
// of the C API (collide_api.cpp). Every function keeps the reference's operation order
// (src/distance.c), so with -ffp-contract=off and correctly rounded division and square root it
// gives the reference's floats. Proxies, simplexes and caches are the public API's structs.
#pragma once

#include "narrowphase.hpp"

namespace b2mi
{

// src/distance.c:13-30
B2MI_HD inline b2Transform sweep_transform( const b2Sweep& sweep, float time )
{
	const float s = 1.0f - time;
	b2Transform xf;
	xf.p = b2Add( b2MulSV( s, sweep.c1 ), b2MulSV( time, sweep.c2 ) );
	b2Rot q;
	q.c = s * sweep.q1.c + time * sweep.q2.c;
	q.s = s * sweep.q1.s + time * sweep.q2.s;
	xf.q = b2NormalizeRot( q );
	xf.p = b2Sub( xf.p, b2RotateVector( xf.q, sweep.localCenter ) );
	return xf;
}

// src/distance.c:109-120
B2MI_HD inline b2DistanceProxy make_proxy( const b2Vec2* points, int count, float radius )
{
	b2DistanceProxy proxy;
	count = count < b2_maxPolygonVertices ? count : b2_maxPolygonVertices;
	for ( int i = 0; i < count; ++i )
		proxy.points[i] = points[i];
	proxy.count = count;
	proxy.radius = radius;
	return proxy;
}

// src/distance.c:132-147: the first point of largest projection
B2MI_HD inline int find_support( const b2DistanceProxy& proxy, b2Vec2 d )
{
	int best = 0;
	float bestValue = b2Dot( proxy.points[0], d );
	for ( int i = 1; i < proxy.count; ++i )
	{
		float value = b2Dot( proxy.points[i], d );
		if ( value > bestValue )
		{
			best = i;
			bestValue = value;
		}
	}
	return best;
}

B2MI_HD inline b2SimplexVertex& simplex_vertex( b2Simplex& s, int i )
{
	return i == 0 ? s.v1 : ( i == 1 ? s.v2 : s.v3 );
}

// the simplex vertex of support points a (on A) and b (on B)
B2MI_HD inline void set_support( b2SimplexVertex& v, const b2DistanceProxy& proxyA, b2Transform xfA, int a, const b2DistanceProxy& proxyB,
								 b2Transform xfB, int b )
{
	v.indexA = a;
	v.indexB = b;
	v.wA = b2TransformPoint( xfA, proxyA.points[a] );
	v.wB = b2TransformPoint( xfB, proxyB.points[b] );
	v.w = b2Sub( v.wB, v.wA );
}

// src/distance.c:149-190: the cached vertices, or the first points when the cache is empty
B2MI_HD inline b2Simplex simplex_from_cache( const b2DistanceCache& cache, const b2DistanceProxy& proxyA, b2Transform xfA,
											 const b2DistanceProxy& proxyB, b2Transform xfB )
{
	b2Simplex s;
	s.count = cache.count;
	for ( int i = 0; i < s.count; ++i )
	{
		b2SimplexVertex& v = simplex_vertex( s, i );
		set_support( v, proxyA, xfA, cache.indexA[i], proxyB, xfB, cache.indexB[i] );
		v.a = -1.0f;
	}
	if ( s.count == 0 )
	{
		set_support( s.v1, proxyA, xfA, 0, proxyB, xfB, 0 );
		s.v1.a = 1.0f;
		s.count = 1;
	}
	return s;
}

// src/distance.c:210-237: from the simplex's closest feature toward the origin
B2MI_HD inline b2Vec2 simplex_search_direction( const b2Simplex& s )
{
	if ( s.count == 1 )
		return b2Neg( s.v1.w );
	b2Vec2 e12 = b2Sub( s.v2.w, s.v1.w );
	float sgn = b2Cross( e12, b2Neg( s.v1.w ) );
	return sgn > 0.0f ? b2LeftPerp( e12 ) : b2RightPerp( e12 );
}

// src/distance.c:122-130
B2MI_HD inline b2Vec2 weight2( float a1, b2Vec2 w1, float a2, b2Vec2 w2 )
{
	return v2( a1 * w1.x + a2 * w2.x, a1 * w1.y + a2 * w2.y );
}

B2MI_HD inline b2Vec2 weight3( float a1, b2Vec2 w1, float a2, b2Vec2 w2, float a3, b2Vec2 w3 )
{
	return v2( a1 * w1.x + a2 * w2.x + a3 * w3.x, a1 * w1.y + a2 * w2.y + a3 * w3.y );
}

// src/distance.c:262-314 (a triangle reports A's point for both)
B2MI_HD inline void simplex_witness_points( b2Vec2& a, b2Vec2& b, const b2Simplex& s )
{
	if ( s.count == 1 )
	{
		a = s.v1.wA;
		b = s.v1.wB;
	}
	else if ( s.count == 2 )
	{
		a = weight2( s.v1.a, s.v1.wA, s.v2.a, s.v2.wA );
		b = weight2( s.v1.a, s.v1.wB, s.v2.a, s.v2.wB );
	}
	else
	{
		a = weight3( s.v1.a, s.v1.wA, s.v2.a, s.v2.wA, s.v3.a, s.v3.wA );
		b = a;
	}
}

// src/distance.c:316-348: the segment's closest point by barycentric coordinates
B2MI_HD inline void solve_simplex2( b2Simplex& s )
{
	b2Vec2 w1 = s.v1.w, w2 = s.v2.w;
	b2Vec2 e12 = b2Sub( w2, w1 );
	float d12_2 = -b2Dot( w1, e12 );
	if ( d12_2 <= 0.0f )
	{
		s.v1.a = 1.0f;
		s.count = 1;
		return;
	}
	float d12_1 = b2Dot( w2, e12 );
	if ( d12_1 <= 0.0f )
	{
		s.v2.a = 1.0f;
		s.count = 1;
		s.v1 = s.v2;
		return;
	}
	float inv = 1.0f / ( d12_1 + d12_2 );
	s.v1.a = d12_1 * inv;
	s.v2.a = d12_2 * inv;
	s.count = 2;
}

// src/distance.c:350-457: the triangle's closest feature (vertex, edge or interior)
B2MI_HD inline void solve_simplex3( b2Simplex& s )
{
	b2Vec2 w1 = s.v1.w, w2 = s.v2.w, w3 = s.v3.w;
	b2Vec2 e12 = b2Sub( w2, w1 );
	float d12_1 = b2Dot( w2, e12 );
	float d12_2 = -b2Dot( w1, e12 );
	b2Vec2 e13 = b2Sub( w3, w1 );
	float d13_1 = b2Dot( w3, e13 );
	float d13_2 = -b2Dot( w1, e13 );
	b2Vec2 e23 = b2Sub( w3, w2 );
	float d23_1 = b2Dot( w3, e23 );
	float d23_2 = -b2Dot( w2, e23 );
	float n123 = b2Cross( e12, e13 );
	float d123_1 = n123 * b2Cross( w2, w3 );
	float d123_2 = n123 * b2Cross( w3, w1 );
	float d123_3 = n123 * b2Cross( w1, w2 );

	if ( d12_2 <= 0.0f && d13_2 <= 0.0f )
	{
		s.v1.a = 1.0f;
		s.count = 1;
		return;
	}
	if ( d12_1 > 0.0f && d12_2 > 0.0f && d123_3 <= 0.0f )
	{
		float inv = 1.0f / ( d12_1 + d12_2 );
		s.v1.a = d12_1 * inv;
		s.v2.a = d12_2 * inv;
		s.count = 2;
		return;
	}
	if ( d13_1 > 0.0f && d13_2 > 0.0f && d123_2 <= 0.0f )
	{
		float inv = 1.0f / ( d13_1 + d13_2 );
		s.v1.a = d13_1 * inv;
		s.v3.a = d13_2 * inv;
		s.count = 2;
		s.v2 = s.v3;
		return;
	}
	if ( d12_1 <= 0.0f && d23_2 <= 0.0f )
	{
		s.v2.a = 1.0f;
		s.count = 1;
		s.v1 = s.v2;
		return;
	}
	if ( d13_1 <= 0.0f && d23_1 <= 0.0f )
	{
		s.v3.a = 1.0f;
		s.count = 1;
		s.v1 = s.v3;
		return;
	}
	if ( d23_1 > 0.0f && d23_2 > 0.0f && d123_1 <= 0.0f )
	{
		float inv = 1.0f / ( d23_1 + d23_2 );
		s.v2.a = d23_1 * inv;
		s.v3.a = d23_2 * inv;
		s.count = 2;
		s.v1 = s.v3;
		return;
	}
	float inv = 1.0f / ( d123_1 + d123_2 + d123_3 );
	s.v1.a = d123_1 * inv;
	s.v2.a = d123_2 * inv;
	s.v3.a = d123_3 * inv;
	s.count = 3;
}

// src/distance.c:459-623: GJK from the cached simplex, at most 20 support calls, stopping on a
// repeated support pair; the cache is updated, and the simplex of each iteration recorded in
// `simplexes` when one is given.
B2MI_HD inline b2DistanceOutput shape_distance( b2DistanceCache& cache, const b2DistanceInput& input, b2Simplex* simplexes = nullptr,
												int simplexCapacity = 0 )
{
	b2DistanceOutput out;
	out.pointA = v2( 0.0f, 0.0f );
	out.pointB = v2( 0.0f, 0.0f );
	const b2DistanceProxy& proxyA = input.proxyA;
	const b2DistanceProxy& proxyB = input.proxyB;
	const b2Transform xfA = input.transformA, xfB = input.transformB;
	b2Simplex s = simplex_from_cache( cache, proxyA, xfA, proxyB, xfB );
	int simplexIndex = 0;
	auto record = [&]() {
		if ( simplexes != nullptr && simplexIndex < simplexCapacity )
			simplexes[simplexIndex++] = s;
	};
	record();
	const int maxIters = 20;
	int saveA[3], saveB[3];
	int iter = 0;
	while ( iter < maxIters )
	{
		const int saveCount = s.count;
		for ( int i = 0; i < saveCount; ++i )
		{
			saveA[i] = simplex_vertex( s, i ).indexA;
			saveB[i] = simplex_vertex( s, i ).indexB;
		}
		if ( s.count == 2 )
			solve_simplex2( s );
		else if ( s.count == 3 )
			solve_simplex3( s );
		if ( s.count == 3 )
			break;
		record();
		b2Vec2 d = simplex_search_direction( s );
		if ( b2Dot( d, d ) < FLT_EPSILON * FLT_EPSILON )
			break;
		b2SimplexVertex& v = simplex_vertex( s, s.count );
		const int a = find_support( proxyA, b2InvRotateVector( xfA.q, b2Neg( d ) ) );
		const int b = find_support( proxyB, b2InvRotateVector( xfB.q, d ) );
		set_support( v, proxyA, xfA, a, proxyB, xfB, b );
		++iter;
		bool duplicate = false;
		for ( int i = 0; i < saveCount; ++i )
			if ( v.indexA == saveA[i] && v.indexB == saveB[i] )
			{
				duplicate = true;
				break;
			}
		if ( duplicate )
			break;
		++s.count;
	}
	record();
	simplex_witness_points( out.pointA, out.pointB, s );
	out.distance = b2Distance( out.pointA, out.pointB );
	out.iterations = iter;
	out.simplexCount = simplexIndex;
	cache.count = (uint16_t)s.count;
	for ( int i = 0; i < s.count; ++i )
	{
		cache.indexA[i] = (uint8_t)simplex_vertex( s, i ).indexA;
		cache.indexB[i] = (uint8_t)simplex_vertex( s, i ).indexB;
	}
	if ( input.useRadii )
	{
		if ( out.distance < FLT_EPSILON )
		{
			b2Vec2 p = v2( 0.5f * ( out.pointA.x + out.pointB.x ), 0.5f * ( out.pointA.y + out.pointB.y ) );
			out.pointA = p;
			out.pointB = p;
			out.distance = 0.0f;
		}
		else
		{
			const float rA = proxyA.radius, rB = proxyB.radius;
			out.distance = maxr( 0.0f, out.distance - rA - rB );
			b2Vec2 normal = d_normalize( b2Sub( out.pointB, out.pointA ) );
			out.pointA = b2Add( out.pointA, v2( rA * normal.x, rA * normal.y ) );
			out.pointB = b2Sub( out.pointB, v2( rB * normal.x, rB * normal.y ) );
		}
	}
	return out;
}

// src/distance.c:787-878: the separating axis of the cached simplex at t1, as two witness
// points, a face of A or a face of B
enum SeparationType : int
{
	kSepPoints = 0,
	kSepFaceA = 1,
	kSepFaceB = 2,
};

struct SeparationFunction
{
	const b2DistanceProxy* proxyA;
	const b2DistanceProxy* proxyB;
	b2Sweep sweepA, sweepB;
	b2Vec2 localPoint;
	b2Vec2 axis;
	int type;
};

B2MI_HD inline SeparationFunction make_separation( const b2DistanceCache& cache, const b2DistanceProxy& proxyA, const b2Sweep& sweepA,
													const b2DistanceProxy& proxyB, const b2Sweep& sweepB, float t1 )
{
	SeparationFunction f;
	f.proxyA = &proxyA;
	f.proxyB = &proxyB;
	f.sweepA = sweepA;
	f.sweepB = sweepB;
	const b2Transform xfA = sweep_transform( sweepA, t1 ), xfB = sweep_transform( sweepB, t1 );
	if ( cache.count == 1 )
	{
		f.type = kSepPoints;
		b2Vec2 pA = b2TransformPoint( xfA, proxyA.points[cache.indexA[0]] );
		b2Vec2 pB = b2TransformPoint( xfB, proxyB.points[cache.indexB[0]] );
		f.axis = d_normalize( b2Sub( pB, pA ) );
		f.localPoint = v2( 0.0f, 0.0f );
		return f;
	}
	if ( cache.indexA[0] == cache.indexA[1] )
	{
		// an edge of B and a point of A
		f.type = kSepFaceB;
		b2Vec2 b1 = proxyB.points[cache.indexB[0]], b2 = proxyB.points[cache.indexB[1]];
		f.axis = d_normalize( b2CrossVS( b2Sub( b2, b1 ), 1.0f ) );
		b2Vec2 normal = b2RotateVector( xfB.q, f.axis );
		f.localPoint = v2( 0.5f * ( b1.x + b2.x ), 0.5f * ( b1.y + b2.y ) );
		b2Vec2 pB = b2TransformPoint( xfB, f.localPoint );
		b2Vec2 pA = b2TransformPoint( xfA, proxyA.points[cache.indexA[0]] );
		if ( b2Dot( b2Sub( pA, pB ), normal ) < 0.0f )
			f.axis = b2Neg( f.axis );
		return f;
	}
	// an edge of A and a point of B
	f.type = kSepFaceA;
	b2Vec2 a1 = proxyA.points[cache.indexA[0]], a2 = proxyA.points[cache.indexA[1]];
	f.axis = d_normalize( b2CrossVS( b2Sub( a2, a1 ), 1.0f ) );
	b2Vec2 normal = b2RotateVector( xfA.q, f.axis );
	f.localPoint = v2( 0.5f * ( a1.x + a2.x ), 0.5f * ( a1.y + a2.y ) );
	b2Vec2 pA = b2TransformPoint( xfA, f.localPoint );
	b2Vec2 pB = b2TransformPoint( xfB, proxyB.points[cache.indexB[0]] );
	if ( b2Dot( b2Sub( pB, pA ), normal ) < 0.0f )
		f.axis = b2Neg( f.axis );
	return f;
}

// src/distance.c:880-946: the deepest points along the axis at time t
B2MI_HD inline float find_min_separation( const SeparationFunction& f, int& indexA, int& indexB, float t )
{
	const b2Transform xfA = sweep_transform( f.sweepA, t ), xfB = sweep_transform( f.sweepB, t );
	if ( f.type == kSepPoints )
	{
		indexA = find_support( *f.proxyA, b2InvRotateVector( xfA.q, f.axis ) );
		indexB = find_support( *f.proxyB, b2InvRotateVector( xfB.q, b2Neg( f.axis ) ) );
		b2Vec2 pA = b2TransformPoint( xfA, f.proxyA->points[indexA] );
		b2Vec2 pB = b2TransformPoint( xfB, f.proxyB->points[indexB] );
		return b2Dot( b2Sub( pB, pA ), f.axis );
	}
	if ( f.type == kSepFaceA )
	{
		b2Vec2 normal = b2RotateVector( xfA.q, f.axis );
		b2Vec2 pA = b2TransformPoint( xfA, f.localPoint );
		indexA = -1;
		indexB = find_support( *f.proxyB, b2InvRotateVector( xfB.q, b2Neg( normal ) ) );
		b2Vec2 pB = b2TransformPoint( xfB, f.proxyB->points[indexB] );
		return b2Dot( b2Sub( pB, pA ), normal );
	}
	b2Vec2 normal = b2RotateVector( xfB.q, f.axis );
	b2Vec2 pB = b2TransformPoint( xfB, f.localPoint );
	indexB = -1;
	indexA = find_support( *f.proxyA, b2InvRotateVector( xfA.q, b2Neg( normal ) ) );
	b2Vec2 pA = b2TransformPoint( xfA, f.proxyA->points[indexA] );
	return b2Dot( b2Sub( pA, pB ), normal );
}

// src/distance.c:948-997: the separation of the given witness points at time t
B2MI_HD inline float evaluate_separation( const SeparationFunction& f, int indexA, int indexB, float t )
{
	const b2Transform xfA = sweep_transform( f.sweepA, t ), xfB = sweep_transform( f.sweepB, t );
	if ( f.type == kSepPoints )
	{
		b2Vec2 pA = b2TransformPoint( xfA, f.proxyA->points[indexA] );
		b2Vec2 pB = b2TransformPoint( xfB, f.proxyB->points[indexB] );
		return b2Dot( b2Sub( pB, pA ), f.axis );
	}
	if ( f.type == kSepFaceA )
	{
		b2Vec2 normal = b2RotateVector( xfA.q, f.axis );
		b2Vec2 pA = b2TransformPoint( xfA, f.localPoint );
		b2Vec2 pB = b2TransformPoint( xfB, f.proxyB->points[indexB] );
		return b2Dot( b2Sub( pB, pA ), normal );
	}
	b2Vec2 normal = b2RotateVector( xfB.q, f.axis );
	b2Vec2 pB = b2TransformPoint( xfB, f.localPoint );
	b2Vec2 pA = b2TransformPoint( xfA, f.proxyA->points[indexA] );
	return b2Dot( b2Sub( pA, pB ), normal );
}

// src/distance.c:999-1237: conservative advancement on successive separating axes, each axis's
// time found by alternating bisection and secant steps (at most 20 axes, 50 root steps each and
// b2_maxPolygonVertices push-backs per axis)
B2MI_HD inline b2TOIOutput time_of_impact( const b2TOIInput& input )
{
	b2TOIOutput out;
	out.state = b2_toiStateUnknown;
	out.t = input.tMax;
	const b2DistanceProxy& proxyA = input.proxyA;
	const b2DistanceProxy& proxyB = input.proxyB;
	const b2Sweep sweepA = input.sweepA, sweepB = input.sweepB;
	const float tMax = input.tMax;
	const float totalRadius = proxyA.radius + proxyB.radius;
	const float target = maxr( K_LINEAR_SLOP, totalRadius - K_LINEAR_SLOP );
	const float tolerance = 0.25f * K_LINEAR_SLOP;
	float t1 = 0.0f;
	const int maxIterations = 20;
	int iter = 0;
	b2DistanceCache cache;
	cache.count = 0;
	for ( int i = 0; i < 3; ++i )
		cache.indexA[i] = cache.indexB[i] = 0;
	b2DistanceInput din;
	din.proxyA = proxyA;
	din.proxyB = proxyB;
	din.useRadii = false;
	for ( ;; )
	{
		din.transformA = sweep_transform( sweepA, t1 );
		din.transformB = sweep_transform( sweepB, t1 );
		const b2DistanceOutput dout = shape_distance( cache, din );
		if ( dout.distance <= 0.0f )
		{
			out.state = b2_toiStateOverlapped;
			out.t = 0.0f;
			break;
		}
		if ( dout.distance < target + tolerance )
		{
			out.state = b2_toiStateHit;
			out.t = t1;
			break;
		}
		const SeparationFunction fcn = make_separation( cache, proxyA, sweepA, proxyB, sweepB, t1 );
		bool done = false;
		float t2 = tMax;
		int pushBackIter = 0;
		for ( ;; )
		{
			int indexA, indexB;
			float s2 = find_min_separation( fcn, indexA, indexB, t2 );
			if ( s2 > target + tolerance )
			{
				out.state = b2_toiStateSeparated;
				out.t = tMax;
				done = true;
				break;
			}
			if ( s2 > target - tolerance )
			{
				t1 = t2;
				break;
			}
			float s1 = evaluate_separation( fcn, indexA, indexB, t1 );
			if ( s1 < target - tolerance )
			{
				out.state = b2_toiStateFailed;
				out.t = t1;
				done = true;
				break;
			}
			if ( s1 <= target + tolerance )
			{
				out.state = b2_toiStateHit;
				out.t = t1;
				done = true;
				break;
			}
			int rootIterCount = 0;
			float a1 = t1, a2 = t2;
			for ( ;; )
			{
				float t;
				if ( rootIterCount & 1 )
					t = a1 + ( target - s1 ) * ( a2 - a1 ) / ( s2 - s1 );
				else
					t = 0.5f * ( a1 + a2 );
				++rootIterCount;
				float s = evaluate_separation( fcn, indexA, indexB, t );
				if ( b2AbsFloat( s - target ) < tolerance )
				{
					t2 = t;
					break;
				}
				if ( s > target )
				{
					a1 = t;
					s1 = s;
				}
				else
				{
					a2 = t;
					s2 = s;
				}
				if ( rootIterCount == 50 )
					break;
			}
			++pushBackIter;
			if ( pushBackIter == b2_maxPolygonVertices )
				break;
		}
		++iter;
		if ( done )
			break;
		if ( iter == maxIterations )
		{
			out.state = b2_toiStateFailed;
			out.t = t1;
			break;
		}
	}
	return out;
}

} // namespace b2mi
