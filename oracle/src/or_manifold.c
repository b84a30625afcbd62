// CPU ORACLE — test infrastructure only (see or_internal.h).
// Narrowphase contact manifolds, restating src/manifold.c:14-712 and the segment distance of
// src/distance.c:32-105 with the same float operation order.
#include "or_internal.h"

#include <float.h>

#define OR_MAKE_ID( A, B ) ( (uint8_t)( A ) << 8 | (uint8_t)( B ) )

// src/manifold.c:14-30
static b2Polygon or_make_capsule( b2Vec2 p1, b2Vec2 p2, float radius )
{
	b2Polygon shape = { 0 };
	shape.vertices[0] = p1;
	shape.vertices[1] = p2;
	shape.centroid = b2Lerp( p1, p2, 0.5f );
	b2Vec2 d = b2Sub( p2, p1 );
	float length = b2Length( d );
	b2Vec2 axis = b2Vec2_zero;
	if ( length >= FLT_EPSILON )
	{
		float inv = 1.0f / length;
		axis = ( b2Vec2 ){ inv * d.x, inv * d.y };
	}
	b2Vec2 normal = b2RightPerp( axis );
	shape.normals[0] = normal;
	shape.normals[1] = b2Neg( normal );
	shape.count = 2;
	shape.radius = radius;
	return shape;
}

// src/distance.c:32-105 (Ericson 5.1.9)
B2_API b2SegmentDistanceResult b2SegmentDistance( b2Vec2 p1, b2Vec2 q1, b2Vec2 p2, b2Vec2 q2 )
{
	b2SegmentDistanceResult result = { 0 };
	b2Vec2 d1 = b2Sub( q1, p1 );
	b2Vec2 d2 = b2Sub( q2, p2 );
	b2Vec2 r = b2Sub( p1, p2 );
	float dd1 = b2Dot( d1, d1 );
	float dd2 = b2Dot( d2, d2 );
	float rd2 = b2Dot( r, d2 );
	float rd1 = b2Dot( r, d1 );
	const float epsSqr = FLT_EPSILON * FLT_EPSILON;

	if ( dd1 < epsSqr || dd2 < epsSqr )
	{
		if ( dd1 >= epsSqr )
		{
			result.fraction1 = b2ClampFloat( -rd1 / dd1, 0.0f, 1.0f );
			result.fraction2 = 0.0f;
		}
		else if ( dd2 >= epsSqr )
		{
			result.fraction1 = 0.0f;
			result.fraction2 = b2ClampFloat( rd2 / dd2, 0.0f, 1.0f );
		}
		else
		{
			result.fraction1 = 0.0f;
			result.fraction2 = 0.0f;
		}
	}
	else
	{
		float d12 = b2Dot( d1, d2 );
		float denom = dd1 * dd2 - d12 * d12;
		float f1 = 0.0f;
		if ( denom != 0.0f )
			f1 = b2ClampFloat( ( d12 * rd2 - rd1 * dd2 ) / denom, 0.0f, 1.0f );
		float f2 = ( d12 * f1 + rd2 ) / dd2;
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
		result.fraction1 = f1;
		result.fraction2 = f2;
	}

	result.closest1 = b2MulAdd( p1, result.fraction1, d1 );
	result.closest2 = b2MulAdd( p2, result.fraction2, d2 );
	result.distanceSquared = b2DistanceSquared( result.closest1, result.closest2 );
	return result;
}

// Fills the world-space parts of a single-point manifold (shared tail of manifold.c:57-68).
static void or_finish_point( b2Manifold* m, b2Transform xfA, b2Transform xfB, b2Vec2 localNormal, b2Vec2 contactPointA,
							 float separation )
{
	m->normal = b2RotateVector( xfA.q, localNormal );
	b2ManifoldPoint* mp = m->points + 0;
	mp->anchorA = b2RotateVector( xfA.q, contactPointA );
	mp->anchorB = b2Add( mp->anchorA, b2Sub( xfA.p, xfB.p ) );
	mp->point = b2Add( xfA.p, mp->anchorA );
	mp->separation = separation;
	mp->id = 0;
	m->pointCount = 1;
}

// src/manifold.c:36-70
B2_API b2Manifold b2CollideCircles( const b2Circle* circleA, b2Transform xfA, const b2Circle* circleB, b2Transform xfB )
{
	b2Manifold manifold = { 0 };
	b2Transform xf = b2InvMulTransforms( xfA, xfB );
	b2Vec2 pointA = circleA->center;
	b2Vec2 pointB = b2TransformPoint( xf, circleB->center );
	float distance;
	b2Vec2 normal = b2GetLengthAndNormalize( &distance, b2Sub( pointB, pointA ) );
	float radiusA = circleA->radius, radiusB = circleB->radius;
	float separation = distance - radiusA - radiusB;
	if ( separation > OR_SPECULATIVE )
		return manifold;
	b2Vec2 cA = b2MulAdd( pointA, radiusA, normal );
	b2Vec2 cB = b2MulAdd( pointB, -radiusB, normal );
	or_finish_point( &manifold, xfA, xfB, normal, b2Lerp( cA, cB, 0.5f ), separation );
	return manifold;
}

// src/manifold.c:73-135
B2_API b2Manifold b2CollideCapsuleAndCircle( const b2Capsule* capsuleA, b2Transform xfA, const b2Circle* circleB,
											 b2Transform xfB )
{
	b2Manifold manifold = { 0 };
	b2Transform xf = b2InvMulTransforms( xfA, xfB );
	b2Vec2 pB = b2TransformPoint( xf, circleB->center );
	b2Vec2 p1 = capsuleA->center1, p2 = capsuleA->center2;
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
	b2Vec2 normal = b2GetLengthAndNormalize( &distance, b2Sub( pB, pA ) );
	float radiusA = capsuleA->radius, radiusB = circleB->radius;
	float separation = distance - radiusA - radiusB;
	if ( separation > OR_SPECULATIVE )
		return manifold;
	b2Vec2 cA = b2MulAdd( pA, radiusA, normal );
	b2Vec2 cB = b2MulAdd( pB, -radiusB, normal );
	or_finish_point( &manifold, xfA, xfB, normal, b2Lerp( cA, cB, 0.5f ), separation );
	return manifold;
}

// src/manifold.c:137-253
B2_API b2Manifold b2CollidePolygonAndCircle( const b2Polygon* polygonA, b2Transform xfA, const b2Circle* circleB,
											 b2Transform xfB )
{
	b2Manifold manifold = { 0 };
	b2Transform xf = b2InvMulTransforms( xfA, xfB );
	b2Vec2 c = b2TransformPoint( xf, circleB->center );
	float radiusA = polygonA->radius, radiusB = circleB->radius;
	float radius = radiusA + radiusB;

	int normalIndex = 0;
	float separation = -FLT_MAX;
	int vertexCount = polygonA->count;
	const b2Vec2* vertices = polygonA->vertices;
	const b2Vec2* normals = polygonA->normals;
	for ( int i = 0; i < vertexCount; ++i )
	{
		float s = b2Dot( normals[i], b2Sub( c, vertices[i] ) );
		if ( s > separation )
		{
			separation = s;
			normalIndex = i;
		}
	}
	if ( separation > radius + OR_SPECULATIVE )
		return manifold;

	int vertIndex1 = normalIndex;
	int vertIndex2 = vertIndex1 + 1 < vertexCount ? vertIndex1 + 1 : 0;
	b2Vec2 v1 = vertices[vertIndex1], v2 = vertices[vertIndex2];
	float u1 = b2Dot( b2Sub( c, v1 ), b2Sub( v2, v1 ) );
	float u2 = b2Dot( b2Sub( c, v2 ), b2Sub( v1, v2 ) );

	if ( ( u1 < 0.0f || u2 < 0.0f ) && separation > FLT_EPSILON )
	{
		// closest to a vertex and safely outside (manifold.c:177-223)
		b2Vec2 v = u1 < 0.0f ? v1 : v2;
		b2Vec2 normal = b2Normalize( b2Sub( c, v ) );
		separation = b2Dot( b2Sub( c, v ), normal );
		if ( separation > radius + OR_SPECULATIVE )
			return manifold;
		b2Vec2 cA = b2MulAdd( v, radiusA, normal );
		b2Vec2 cB = b2MulSub( c, radiusB, normal );
		or_finish_point( &manifold, xfA, xfB, normal, b2Lerp( cA, cB, 0.5f ), b2Dot( b2Sub( cB, cA ), normal ) );
	}
	else
	{
		// center between v1 and v2, possibly inside (manifold.c:224-250)
		b2Vec2 normal = normals[normalIndex];
		b2Vec2 cA = b2MulAdd( c, radiusA - b2Dot( b2Sub( c, v1 ), normal ), normal );
		b2Vec2 cB = b2MulSub( c, radiusB, normal );
		or_finish_point( &manifold, xfA, xfB, normal, b2Lerp( cA, cB, 0.5f ), separation - radius );
	}
	return manifold;
}

// src/manifold.c:276-411
static b2Manifold or_clip_polygons( const b2Polygon* polyA, const b2Polygon* polyB, int edgeA, int edgeB, bool flip )
{
	b2Manifold manifold = { 0 };
	const b2Polygon *poly1, *poly2;
	int i11, i12, i21, i22;
	if ( flip )
	{
		poly1 = polyB;
		poly2 = polyA;
		i11 = edgeB;
		i12 = edgeB + 1 < polyB->count ? edgeB + 1 : 0;
		i21 = edgeA;
		i22 = edgeA + 1 < polyA->count ? edgeA + 1 : 0;
	}
	else
	{
		poly1 = polyA;
		poly2 = polyB;
		i11 = edgeA;
		i12 = edgeA + 1 < polyA->count ? edgeA + 1 : 0;
		i21 = edgeB;
		i22 = edgeB + 1 < polyB->count ? edgeB + 1 : 0;
	}

	b2Vec2 normal = poly1->normals[i11];
	b2Vec2 v11 = poly1->vertices[i11], v12 = poly1->vertices[i12];
	b2Vec2 v21 = poly2->vertices[i21], v22 = poly2->vertices[i22];
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
		manifold.normal = normal;
		manifold.points[0].anchorA = vLower;
		manifold.points[0].separation = separationLower - radius;
		manifold.points[0].id = OR_MAKE_ID( i11, i22 );
		manifold.points[1].anchorA = vUpper;
		manifold.points[1].separation = separationUpper - radius;
		manifold.points[1].id = OR_MAKE_ID( i12, i21 );
		manifold.pointCount = 2;
	}
	else
	{
		manifold.normal = b2Neg( normal );
		manifold.points[0].anchorA = vUpper;
		manifold.points[0].separation = separationUpper - radius;
		manifold.points[0].id = OR_MAKE_ID( i21, i12 );
		manifold.points[1].anchorA = vLower;
		manifold.points[1].separation = separationLower - radius;
		manifold.points[1].id = OR_MAKE_ID( i22, i11 );
		manifold.pointCount = 2;
	}
	return manifold;
}

// src/manifold.c:414-450
static float or_find_max_separation( int* edgeIndex, const b2Polygon* poly1, const b2Polygon* poly2 )
{
	int count1 = poly1->count, count2 = poly2->count;
	int bestIndex = 0;
	float maxSeparation = -FLT_MAX;
	for ( int i = 0; i < count1; ++i )
	{
		b2Vec2 n = poly1->normals[i];
		b2Vec2 v1 = poly1->vertices[i];
		float si = FLT_MAX;
		for ( int j = 0; j < count2; ++j )
		{
			float sij = b2Dot( n, b2Sub( poly2->vertices[j], v1 ) );
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

// single vertex-vertex point of manifold.c:571-663
static bool or_vertex_point( b2Manifold* m, b2Vec2 vA, b2Vec2 vB, float distanceSquared, float rA, float rB, float radius,
							 int idA, int idB )
{
	b2Vec2 normal = b2Sub( vB, vA );
	float distance = sqrtf( distanceSquared );
	if ( distance > OR_SPECULATIVE + radius )
		return false;
	float invDistance = 1.0f / distance;
	normal.x *= invDistance;
	normal.y *= invDistance;
	b2Vec2 c1 = b2MulAdd( vA, rA, normal );
	b2Vec2 c2 = b2MulAdd( vB, -rB, normal );
	m->normal = normal;
	m->points[0].anchorA = b2Lerp( c1, c2, 0.5f );
	m->points[0].separation = distance - radius;
	m->points[0].id = OR_MAKE_ID( idA, idB );
	m->pointCount = 1;
	return true;
}

// src/manifold.c:470-700
B2_API b2Manifold b2CollidePolygons( const b2Polygon* polygonA, b2Transform xfA, const b2Polygon* polygonB, b2Transform xfB )
{
	b2Vec2 origin = polygonA->vertices[0];
	b2Transform sfA = { b2Add( xfA.p, b2RotateVector( xfA.q, origin ) ), xfA.q };
	b2Transform xf = b2InvMulTransforms( sfA, xfB );

	b2Polygon localPolyA;
	localPolyA.count = polygonA->count;
	localPolyA.radius = polygonA->radius;
	localPolyA.vertices[0] = b2Vec2_zero;
	localPolyA.normals[0] = polygonA->normals[0];
	for ( int i = 1; i < localPolyA.count; ++i )
	{
		localPolyA.vertices[i] = b2Sub( polygonA->vertices[i], origin );
		localPolyA.normals[i] = polygonA->normals[i];
	}

	b2Polygon localPolyB;
	localPolyB.count = polygonB->count;
	localPolyB.radius = polygonB->radius;
	for ( int i = 0; i < localPolyB.count; ++i )
	{
		localPolyB.vertices[i] = b2TransformPoint( xf, polygonB->vertices[i] );
		localPolyB.normals[i] = b2RotateVector( xf.q, polygonB->normals[i] );
	}

	int edgeA = 0;
	float separationA = or_find_max_separation( &edgeA, &localPolyA, &localPolyB );
	int edgeB = 0;
	float separationB = or_find_max_separation( &edgeB, &localPolyB, &localPolyA );
	float radius = localPolyA.radius + localPolyB.radius;

	if ( separationA > OR_SPECULATIVE + radius || separationB > OR_SPECULATIVE + radius )
		return ( b2Manifold ){ 0 };

	bool flip;
	if ( separationA >= separationB )
	{
		flip = false;
		b2Vec2 searchDirection = localPolyA.normals[edgeA];
		edgeB = 0;
		float minDot = FLT_MAX;
		for ( int i = 0; i < localPolyB.count; ++i )
		{
			float dot = b2Dot( searchDirection, localPolyB.normals[i] );
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
		b2Vec2 searchDirection = localPolyB.normals[edgeB];
		edgeA = 0;
		float minDot = FLT_MAX;
		for ( int i = 0; i < localPolyA.count; ++i )
		{
			float dot = b2Dot( searchDirection, localPolyA.normals[i] );
			if ( dot < minDot )
			{
				minDot = dot;
				edgeA = i;
			}
		}
	}

	b2Manifold manifold = { 0 };
	if ( separationA > 0.1f * OR_LINEAR_SLOP || separationB > 0.1f * OR_LINEAR_SLOP )
	{
		int i11 = edgeA;
		int i12 = edgeA + 1 < localPolyA.count ? edgeA + 1 : 0;
		int i21 = edgeB;
		int i22 = edgeB + 1 < localPolyB.count ? edgeB + 1 : 0;
		b2Vec2 v11 = localPolyA.vertices[i11], v12 = localPolyA.vertices[i12];
		b2Vec2 v21 = localPolyB.vertices[i21], v22 = localPolyB.vertices[i22];
		b2SegmentDistanceResult result = b2SegmentDistance( v11, v12, v21, v22 );
		float rA = localPolyA.radius, rB = localPolyB.radius;

		if ( result.fraction1 == 0.0f && result.fraction2 == 0.0f )
		{
			if ( !or_vertex_point( &manifold, v11, v21, result.distanceSquared, rA, rB, radius, i11, i21 ) )
				return manifold;
		}
		else if ( result.fraction1 == 0.0f && result.fraction2 == 1.0f )
		{
			if ( !or_vertex_point( &manifold, v11, v22, result.distanceSquared, rA, rB, radius, i11, i22 ) )
				return manifold;
		}
		else if ( result.fraction1 == 1.0f && result.fraction2 == 0.0f )
		{
			if ( !or_vertex_point( &manifold, v12, v21, result.distanceSquared, rA, rB, radius, i12, i21 ) )
				return manifold;
		}
		else if ( result.fraction1 == 1.0f && result.fraction2 == 1.0f )
		{
			if ( !or_vertex_point( &manifold, v12, v22, result.distanceSquared, rA, rB, radius, i12, i22 ) )
				return manifold;
		}
		else
		{
			manifold = or_clip_polygons( &localPolyA, &localPolyB, edgeA, edgeB, flip );
		}
	}
	else
	{
		manifold = or_clip_polygons( &localPolyA, &localPolyB, edgeA, edgeB, flip );
	}

	if ( manifold.pointCount > 0 )
	{
		manifold.normal = b2RotateVector( xfA.q, manifold.normal );
		for ( int i = 0; i < manifold.pointCount; ++i )
		{
			b2ManifoldPoint* mp = manifold.points + i;
			mp->anchorA = b2RotateVector( xfA.q, b2Add( mp->anchorA, origin ) );
			mp->anchorB = b2Add( mp->anchorA, b2Sub( xfA.p, xfB.p ) );
			mp->point = b2Add( xfA.p, mp->anchorA );
		}
	}
	return manifold;
}

// src/manifold.c:255-260
B2_API b2Manifold b2CollideCapsules( const b2Capsule* capsuleA, b2Transform xfA, const b2Capsule* capsuleB, b2Transform xfB )
{
	b2Polygon polyA = or_make_capsule( capsuleA->center1, capsuleA->center2, capsuleA->radius );
	b2Polygon polyB = or_make_capsule( capsuleB->center1, capsuleB->center2, capsuleB->radius );
	return b2CollidePolygons( &polyA, xfA, &polyB, xfB );
}

// src/manifold.c:262-267
B2_API b2Manifold b2CollideSegmentAndCapsule( const b2Segment* segmentA, b2Transform xfA, const b2Capsule* capsuleB,
											  b2Transform xfB )
{
	b2Polygon polyA = or_make_capsule( segmentA->point1, segmentA->point2, 0.0f );
	b2Polygon polyB = or_make_capsule( capsuleB->center1, capsuleB->center2, capsuleB->radius );
	return b2CollidePolygons( &polyA, xfA, &polyB, xfB );
}

// src/manifold.c:269-273
B2_API b2Manifold b2CollidePolygonAndCapsule( const b2Polygon* polygonA, b2Transform xfA, const b2Capsule* capsuleB,
											  b2Transform xfB )
{
	b2Polygon polyB = or_make_capsule( capsuleB->center1, capsuleB->center2, capsuleB->radius );
	return b2CollidePolygons( polygonA, xfA, &polyB, xfB );
}

// src/manifold.c:702-706
B2_API b2Manifold b2CollideSegmentAndCircle( const b2Segment* segmentA, b2Transform xfA, const b2Circle* circleB,
											 b2Transform xfB )
{
	b2Capsule capsuleA = { segmentA->point1, segmentA->point2, 0.0f };
	return b2CollideCapsuleAndCircle( &capsuleA, xfA, circleB, xfB );
}

// src/manifold.c:708-712
B2_API b2Manifold b2CollideSegmentAndPolygon( const b2Segment* segmentA, b2Transform xfA, const b2Polygon* polygonB,
											  b2Transform xfB )
{
	b2Polygon polygonA = or_make_capsule( segmentA->point1, segmentA->point2, 0.0f );
	return b2CollidePolygons( &polygonA, xfA, polygonB, xfB );
}

// Dispatch table of src/contact.c:150-183: (typeA, typeB) -> manifold, primary when typeA
// is the first argument of the registered function.
bool or_has_manifold_fcn( int typeA, int typeB, bool* primary )
{
	// circle=0 capsule=1 segment=2 polygon=3
	static const signed char table[4][4] = {
		// B:   circle capsule segment polygon
		/* circle  */ { 1, -1, -1, -1 },
		/* capsule */ { 1, 1, -1, -1 },
		/* segment */ { 1, 1, 0, 1 },
		/* polygon */ { 1, 1, -1, 1 },
	};
	if ( typeA < 0 || typeA > 3 || typeB < 0 || typeB > 3 )
		return false;
	int v = table[typeA][typeB];
	if ( v == 0 )
		return false;
	*primary = v > 0;
	return true;
}

b2Manifold or_collide_shapes( int typeA, const void* geomA, b2Transform xfA, int typeB, const void* geomB, b2Transform xfB )
{
	switch ( typeA * 4 + typeB )
	{
		case b2_circleShape * 4 + b2_circleShape:
			return b2CollideCircles( geomA, xfA, geomB, xfB );
		case b2_capsuleShape * 4 + b2_circleShape:
			return b2CollideCapsuleAndCircle( geomA, xfA, geomB, xfB );
		case b2_capsuleShape * 4 + b2_capsuleShape:
			return b2CollideCapsules( geomA, xfA, geomB, xfB );
		case b2_polygonShape * 4 + b2_circleShape:
			return b2CollidePolygonAndCircle( geomA, xfA, geomB, xfB );
		case b2_polygonShape * 4 + b2_capsuleShape:
			return b2CollidePolygonAndCapsule( geomA, xfA, geomB, xfB );
		case b2_polygonShape * 4 + b2_polygonShape:
			return b2CollidePolygons( geomA, xfA, geomB, xfB );
		case b2_segmentShape * 4 + b2_circleShape:
			return b2CollideSegmentAndCircle( geomA, xfA, geomB, xfB );
		case b2_segmentShape * 4 + b2_capsuleShape:
			return b2CollideSegmentAndCapsule( geomA, xfA, geomB, xfB );
		case b2_segmentShape * 4 + b2_polygonShape:
			return b2CollideSegmentAndPolygon( geomA, xfA, geomB, xfB );
		default:
		{
			b2Manifold empty = { 0 };
			return empty;
		}
	}
}
