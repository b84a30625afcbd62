// CPU ORACLE — test infrastructure only (see or_internal.h).
//
// GJK distance and time of impact: a restatement of src/distance.c:13-30 (sweep transform),
// 105-623 (proxy, simplex solver, b2ShapeDistance) and 785-1237 (separation functions and
// b2TimeOfImpact), op for op in float, exported under the reference's names.
#include "or_internal.h"

#include <float.h>
#include <math.h>

// src/distance.c:13-30
B2_API b2Transform b2GetSweepTransform( const b2Sweep* sweep, float time )
{
	b2Transform xf;
	xf.p = b2Add( b2MulSV( 1.0f - time, sweep->c1 ), b2MulSV( time, sweep->c2 ) );
	b2Rot q = { ( 1.0f - time ) * sweep->q1.c + time * sweep->q2.c, ( 1.0f - time ) * sweep->q1.s + time * sweep->q2.s };
	xf.q = b2NormalizeRot( q );
	xf.p = b2Sub( xf.p, b2RotateVector( xf.q, sweep->localCenter ) );
	return xf;
}

// src/distance.c:109-120
B2_API b2DistanceProxy b2MakeProxy( const b2Vec2* vertices, int32_t count, float radius )
{
	b2DistanceProxy proxy;
	if ( count > b2_maxPolygonVertices )
		count = b2_maxPolygonVertices;
	for ( int i = 0; i < count; ++i )
		proxy.points[i] = vertices[i];
	proxy.count = count;
	proxy.radius = radius;
	return proxy;
}

// src/distance.c:132-147
static int or_support( const b2DistanceProxy* proxy, b2Vec2 d )
{
	int best = 0;
	float bestValue = b2Dot( proxy->points[0], d );
	for ( int i = 1; i < proxy->count; ++i )
	{
		float value = b2Dot( proxy->points[i], d );
		if ( value > bestValue )
		{
			best = i;
			bestValue = value;
		}
	}
	return best;
}

static void or_simplex_point( b2SimplexVertex* v, const b2DistanceProxy* pa, b2Transform xa, int ia, const b2DistanceProxy* pb,
							  b2Transform xb, int ib )
{
	v->indexA = ia;
	v->indexB = ib;
	v->wA = b2TransformPoint( xa, pa->points[ia] );
	v->wB = b2TransformPoint( xb, pb->points[ib] );
	v->w = b2Sub( v->wB, v->wA );
}

// src/distance.c:316-348
static void or_solve2( b2Simplex* s )
{
	b2Vec2 w1 = s->v1.w, w2 = s->v2.w;
	b2Vec2 e12 = b2Sub( w2, w1 );
	float d12_2 = -b2Dot( w1, e12 );
	if ( d12_2 <= 0.0f )
	{
		s->v1.a = 1.0f;
		s->count = 1;
		return;
	}
	float d12_1 = b2Dot( w2, e12 );
	if ( d12_1 <= 0.0f )
	{
		s->v2.a = 1.0f;
		s->count = 1;
		s->v1 = s->v2;
		return;
	}
	float inv_d12 = 1.0f / ( d12_1 + d12_2 );
	s->v1.a = d12_1 * inv_d12;
	s->v2.a = d12_2 * inv_d12;
	s->count = 2;
}

// src/distance.c:350-457
static void or_solve3( b2Simplex* s )
{
	b2Vec2 w1 = s->v1.w, w2 = s->v2.w, w3 = s->v3.w;
	b2Vec2 e12 = b2Sub( w2, w1 );
	float d12_1 = b2Dot( w2, e12 ), d12_2 = -b2Dot( w1, e12 );
	b2Vec2 e13 = b2Sub( w3, w1 );
	float d13_1 = b2Dot( w3, e13 ), d13_2 = -b2Dot( w1, e13 );
	b2Vec2 e23 = b2Sub( w3, w2 );
	float d23_1 = b2Dot( w3, e23 ), d23_2 = -b2Dot( w2, e23 );
	float n123 = b2Cross( e12, e13 );
	float d123_1 = n123 * b2Cross( w2, w3 );
	float d123_2 = n123 * b2Cross( w3, w1 );
	float d123_3 = n123 * b2Cross( w1, w2 );

	if ( d12_2 <= 0.0f && d13_2 <= 0.0f )
	{
		s->v1.a = 1.0f;
		s->count = 1;
	}
	else if ( d12_1 > 0.0f && d12_2 > 0.0f && d123_3 <= 0.0f )
	{
		float inv = 1.0f / ( d12_1 + d12_2 );
		s->v1.a = d12_1 * inv;
		s->v2.a = d12_2 * inv;
		s->count = 2;
	}
	else if ( d13_1 > 0.0f && d13_2 > 0.0f && d123_2 <= 0.0f )
	{
		float inv = 1.0f / ( d13_1 + d13_2 );
		s->v1.a = d13_1 * inv;
		s->v3.a = d13_2 * inv;
		s->count = 2;
		s->v2 = s->v3;
	}
	else if ( d12_1 <= 0.0f && d23_2 <= 0.0f )
	{
		s->v2.a = 1.0f;
		s->count = 1;
		s->v1 = s->v2;
	}
	else if ( d13_1 <= 0.0f && d23_1 <= 0.0f )
	{
		s->v3.a = 1.0f;
		s->count = 1;
		s->v1 = s->v3;
	}
	else if ( d23_1 > 0.0f && d23_2 > 0.0f && d123_1 <= 0.0f )
	{
		float inv = 1.0f / ( d23_1 + d23_2 );
		s->v2.a = d23_1 * inv;
		s->v3.a = d23_2 * inv;
		s->count = 2;
		s->v1 = s->v3;
	}
	else
	{
		float inv = 1.0f / ( d123_1 + d123_2 + d123_3 );
		s->v1.a = d123_1 * inv;
		s->v2.a = d123_2 * inv;
		s->v3.a = d123_3 * inv;
		s->count = 3;
	}
}

// src/distance.c:459-623
B2_API b2DistanceOutput b2ShapeDistance( b2DistanceCache* cache, const b2DistanceInput* input, b2Simplex* simplexes, int simplexCapacity )
{
	b2DistanceOutput output = { 0 };
	const b2DistanceProxy* pa = &input->proxyA;
	const b2DistanceProxy* pb = &input->proxyB;
	b2Transform xa = input->transformA, xb = input->transformB;

	// b2MakeSimplexFromCache (src/distance.c:149-190)
	b2Simplex s;
	b2SimplexVertex* v[3] = { &s.v1, &s.v2, &s.v3 };
	s.count = cache->count;
	for ( int i = 0; i < s.count; ++i )
	{
		or_simplex_point( v[i], pa, xa, cache->indexA[i], pb, xb, cache->indexB[i] );
		v[i]->a = -1.0f;
	}
	if ( s.count == 0 )
	{
		or_simplex_point( v[0], pa, xa, 0, pb, xb, 0 );
		v[0]->a = 1.0f;
		s.count = 1;
	}

	int recorded = 0;
	if ( simplexes != NULL && recorded < simplexCapacity )
		simplexes[recorded++] = s;

	int saveA[3], saveB[3];
	int iter = 0;
	while ( iter < 20 )
	{
		int saveCount = s.count;
		for ( int i = 0; i < saveCount; ++i )
		{
			saveA[i] = v[i]->indexA;
			saveB[i] = v[i]->indexB;
		}
		if ( s.count == 2 )
			or_solve2( &s );
		else if ( s.count == 3 )
			or_solve3( &s );
		if ( s.count == 3 )
			break;
		if ( simplexes != NULL && recorded < simplexCapacity )
			simplexes[recorded++] = s;

		// b2ComputeSimplexSearchDirection (src/distance.c:210-237)
		b2Vec2 d;
		if ( s.count == 1 )
			d = b2Neg( s.v1.w );
		else
		{
			b2Vec2 e12 = b2Sub( s.v2.w, s.v1.w );
			d = b2Cross( e12, b2Neg( s.v1.w ) ) > 0.0f ? b2LeftPerp( e12 ) : b2RightPerp( e12 );
		}
		if ( b2Dot( d, d ) < FLT_EPSILON * FLT_EPSILON )
			break;

		b2SimplexVertex* nv = v[s.count];
		int ia = or_support( pa, b2InvRotateVector( xa.q, b2Neg( d ) ) );
		int ib = or_support( pb, b2InvRotateVector( xb.q, d ) );
		or_simplex_point( nv, pa, xa, ia, pb, xb, ib );
		++iter;

		bool duplicate = false;
		for ( int i = 0; i < saveCount && !duplicate; ++i )
			duplicate = nv->indexA == saveA[i] && nv->indexB == saveB[i];
		if ( duplicate )
			break;
		++s.count;
	}
	if ( simplexes != NULL && recorded < simplexCapacity )
		simplexes[recorded++] = s;

	// b2ComputeSimplexWitnessPoints (src/distance.c:262-314)
	if ( s.count == 1 )
	{
		output.pointA = s.v1.wA;
		output.pointB = s.v1.wB;
	}
	else if ( s.count == 2 )
	{
		output.pointA = ( b2Vec2 ){ s.v1.a * s.v1.wA.x + s.v2.a * s.v2.wA.x, s.v1.a * s.v1.wA.y + s.v2.a * s.v2.wA.y };
		output.pointB = ( b2Vec2 ){ s.v1.a * s.v1.wB.x + s.v2.a * s.v2.wB.x, s.v1.a * s.v1.wB.y + s.v2.a * s.v2.wB.y };
	}
	else
	{
		output.pointA = ( b2Vec2 ){ s.v1.a * s.v1.wA.x + s.v2.a * s.v2.wA.x + s.v3.a * s.v3.wA.x,
									s.v1.a * s.v1.wA.y + s.v2.a * s.v2.wA.y + s.v3.a * s.v3.wA.y };
		output.pointB = output.pointA;
	}
	output.distance = b2Distance( output.pointA, output.pointB );
	output.iterations = iter;
	output.simplexCount = recorded;

	// b2MakeSimplexCache (src/distance.c:192-203)
	cache->count = (uint16_t)s.count;
	for ( int i = 0; i < s.count; ++i )
	{
		cache->indexA[i] = (uint8_t)v[i]->indexA;
		cache->indexB[i] = (uint8_t)v[i]->indexB;
	}

	if ( input->useRadii )
	{
		if ( output.distance < FLT_EPSILON )
		{
			b2Vec2 p = { 0.5f * ( output.pointA.x + output.pointB.x ), 0.5f * ( output.pointA.y + output.pointB.y ) };
			output.pointA = p;
			output.pointB = p;
			output.distance = 0.0f;
		}
		else
		{
			float rA = pa->radius, rB = pb->radius;
			output.distance = b2MaxFloat( 0.0f, output.distance - rA - rB );
			b2Vec2 n = b2Normalize( b2Sub( output.pointB, output.pointA ) );
			output.pointA = b2Add( output.pointA, ( b2Vec2 ){ rA * n.x, rA * n.y } );
			output.pointB = b2Sub( output.pointB, ( b2Vec2 ){ rB * n.x, rB * n.y } );
		}
	}
	return output;
}

// src/distance.c:787-802
typedef struct orSeparation
{
	const b2DistanceProxy* pa;
	const b2DistanceProxy* pb;
	b2Sweep sa, sb;
	b2Vec2 localPoint;
	b2Vec2 axis;
	int type; // 0 points, 1 face of A, 2 face of B
} orSeparation;

// src/distance.c:804-878
static orSeparation or_make_separation( const b2DistanceCache* cache, const b2DistanceProxy* pa, const b2Sweep* sa,
										const b2DistanceProxy* pb, const b2Sweep* sb, float t1 )
{
	orSeparation f;
	f.pa = pa;
	f.pb = pb;
	f.sa = *sa;
	f.sb = *sb;
	b2Transform xa = b2GetSweepTransform( sa, t1 ), xb = b2GetSweepTransform( sb, t1 );
	if ( cache->count == 1 )
	{
		f.type = 0;
		b2Vec2 pA = b2TransformPoint( xa, pa->points[cache->indexA[0]] );
		b2Vec2 pB = b2TransformPoint( xb, pb->points[cache->indexB[0]] );
		f.axis = b2Normalize( b2Sub( pB, pA ) );
		f.localPoint = b2Vec2_zero;
		return f;
	}
	if ( cache->indexA[0] == cache->indexA[1] )
	{
		f.type = 2;
		b2Vec2 b1 = pb->points[cache->indexB[0]], b2 = pb->points[cache->indexB[1]];
		f.axis = b2Normalize( b2CrossVS( b2Sub( b2, b1 ), 1.0f ) );
		b2Vec2 normal = b2RotateVector( xb.q, f.axis );
		f.localPoint = ( b2Vec2 ){ 0.5f * ( b1.x + b2.x ), 0.5f * ( b1.y + b2.y ) };
		b2Vec2 pB = b2TransformPoint( xb, f.localPoint );
		b2Vec2 pA = b2TransformPoint( xa, pa->points[cache->indexA[0]] );
		if ( b2Dot( b2Sub( pA, pB ), normal ) < 0.0f )
			f.axis = b2Neg( f.axis );
		return f;
	}
	f.type = 1;
	b2Vec2 a1 = pa->points[cache->indexA[0]], a2 = pa->points[cache->indexA[1]];
	f.axis = b2Normalize( b2CrossVS( b2Sub( a2, a1 ), 1.0f ) );
	b2Vec2 normal = b2RotateVector( xa.q, f.axis );
	f.localPoint = ( b2Vec2 ){ 0.5f * ( a1.x + a2.x ), 0.5f * ( a1.y + a2.y ) };
	b2Vec2 pA = b2TransformPoint( xa, f.localPoint );
	b2Vec2 pB = b2TransformPoint( xb, pb->points[cache->indexB[0]] );
	if ( b2Dot( b2Sub( pB, pA ), normal ) < 0.0f )
		f.axis = b2Neg( f.axis );
	return f;
}

// src/distance.c:880-997: with `find` the deepest points are searched (b2FindMinSeparation),
// otherwise the given ones are evaluated (b2EvaluateSeparation)
static float or_separation( const orSeparation* f, bool find, int* ia, int* ib, float t )
{
	b2Transform xa = b2GetSweepTransform( &f->sa, t ), xb = b2GetSweepTransform( &f->sb, t );
	if ( f->type == 0 )
	{
		if ( find )
		{
			*ia = or_support( f->pa, b2InvRotateVector( xa.q, f->axis ) );
			*ib = or_support( f->pb, b2InvRotateVector( xb.q, b2Neg( f->axis ) ) );
		}
		b2Vec2 pA = b2TransformPoint( xa, f->pa->points[*ia] );
		b2Vec2 pB = b2TransformPoint( xb, f->pb->points[*ib] );
		return b2Dot( b2Sub( pB, pA ), f->axis );
	}
	if ( f->type == 1 )
	{
		b2Vec2 normal = b2RotateVector( xa.q, f->axis );
		b2Vec2 pA = b2TransformPoint( xa, f->localPoint );
		if ( find )
		{
			*ia = -1;
			*ib = or_support( f->pb, b2InvRotateVector( xb.q, b2Neg( normal ) ) );
		}
		b2Vec2 pB = b2TransformPoint( xb, f->pb->points[*ib] );
		return b2Dot( b2Sub( pB, pA ), normal );
	}
	b2Vec2 normal = b2RotateVector( xb.q, f->axis );
	b2Vec2 pB = b2TransformPoint( xb, f->localPoint );
	if ( find )
	{
		*ib = -1;
		*ia = or_support( f->pa, b2InvRotateVector( xa.q, b2Neg( normal ) ) );
	}
	b2Vec2 pA = b2TransformPoint( xa, f->pa->points[*ia] );
	return b2Dot( b2Sub( pA, pB ), normal );
}

// src/distance.c:999-1237
B2_API b2TOIOutput b2TimeOfImpact( const b2TOIInput* input )
{
	b2TOIOutput output = { b2_toiStateUnknown, input->tMax };
	const b2DistanceProxy* pa = &input->proxyA;
	const b2DistanceProxy* pb = &input->proxyB;
	b2Sweep sa = input->sweepA, sb = input->sweepB;
	float tMax = input->tMax;
	float target = b2MaxFloat( OR_LINEAR_SLOP, pa->radius + pb->radius - OR_LINEAR_SLOP );
	float tolerance = 0.25f * OR_LINEAR_SLOP;

	float t1 = 0.0f;
	int iter = 0;
	b2DistanceCache cache = { 0 };
	b2DistanceInput din;
	din.proxyA = *pa;
	din.proxyB = *pb;
	din.useRadii = false;
	for ( ;; )
	{
		din.transformA = b2GetSweepTransform( &sa, t1 );
		din.transformB = b2GetSweepTransform( &sb, t1 );
		b2DistanceOutput dout = b2ShapeDistance( &cache, &din, NULL, 0 );
		if ( dout.distance <= 0.0f )
		{
			output.state = b2_toiStateOverlapped;
			output.t = 0.0f;
			break;
		}
		if ( dout.distance < target + tolerance )
		{
			output.state = b2_toiStateHit;
			output.t = t1;
			break;
		}

		orSeparation fcn = or_make_separation( &cache, pa, &sa, pb, &sb, t1 );
		bool done = false;
		float t2 = tMax;
		for ( int pushBack = 0; pushBack < b2_maxPolygonVertices; )
		{
			int ia = -1, ib = -1;
			float s2 = or_separation( &fcn, true, &ia, &ib, t2 );
			if ( s2 > target + tolerance )
			{
				output.state = b2_toiStateSeparated;
				output.t = tMax;
				done = true;
				break;
			}
			if ( s2 > target - tolerance )
			{
				t1 = t2;
				break;
			}
			float s1 = or_separation( &fcn, false, &ia, &ib, t1 );
			if ( s1 < target - tolerance )
			{
				output.state = b2_toiStateFailed;
				output.t = t1;
				done = true;
				break;
			}
			if ( s1 <= target + tolerance )
			{
				output.state = b2_toiStateHit;
				output.t = t1;
				done = true;
				break;
			}
			// bisection on even steps, secant on odd ones, at most 50
			float a1 = t1, a2 = t2;
			for ( int root = 0; root < 50; )
			{
				float t = ( root & 1 ) ? a1 + ( target - s1 ) * ( a2 - a1 ) / ( s2 - s1 ) : 0.5f * ( a1 + a2 );
				++root;
				float s = or_separation( &fcn, false, &ia, &ib, t );
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
			}
			++pushBack;
		}
		++iter;
		if ( done )
			break;
		if ( iter == 20 )
		{
			output.state = b2_toiStateFailed;
			output.t = t1;
			break;
		}
	}
	return output;
}
