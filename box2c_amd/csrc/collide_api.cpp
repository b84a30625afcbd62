// The reference's public narrowphase, distance and time-of-impact functions
// (include/box2d/collision.h, src/manifold.c, src/distance.c) exported from the host compile of
// the same code that k_collide and k_continuous run on the GPU (narrowphase.hpp, toi.hpp). Returned manifolds carry the points, anchors,
// separations and feature ids; impulses, normal velocity and the persisted flag are zero,
// as the reference's functions leave them.
#include "joints.hpp"
#include "narrowphase.hpp"
#include "toi.hpp"

#include "box2d/collision.h"

#include <string.h>

using namespace b2mi;

namespace
{

Poly to_poly( const b2Polygon* p )
{
	Poly q;
	q.count = p->count;
	q.radius = p->radius;
	for ( int i = 0; i < 8; ++i )
	{
		q.v[i] = i < p->count ? p->vertices[i] : v2( 0.0f, 0.0f );
		q.n[i] = i < p->count ? p->normals[i] : v2( 0.0f, 0.0f );
	}
	return q;
}

Poly capsule_poly( b2Vec2 p1, b2Vec2 p2, float radius )
{
	Poly q;
	make_capsule( q, p1, p2, radius );
	return q;
}

b2Manifold to_public( const Manifold& m )
{
	b2Manifold r;
	memset( &r, 0, sizeof( r ) );
	r.normal = m.normal;
	r.pointCount = m.pointCount;
	for ( int i = 0; i < m.pointCount; ++i )
	{
		b2ManifoldPoint& p = r.points[i];
		p.point = m.point[i];
		p.anchorA = m.anchorA[i];
		p.anchorB = m.anchorB[i];
		p.separation = m.separation[i];
		p.id = m.id[i];
	}
	return r;
}

b2Manifold polygons( const Poly& a, b2Transform xfA, const Poly& b, b2Transform xfB )
{
	Manifold m;
	mf_clear( m );
	collide_polygons( m, a, xfA, b, xfB );
	return to_public( m );
}

} // namespace

extern "C" {

// src/distance.c:32-105
B2_API b2SegmentDistanceResult b2SegmentDistance( b2Vec2 p1, b2Vec2 q1, b2Vec2 p2, b2Vec2 q2 )
{
	b2SegmentDistanceResult r;
	segment_distance( p1, q1, p2, q2, &r.fraction1, &r.fraction2, &r.distanceSquared );
	r.closest1 = b2MulAdd( p1, r.fraction1, b2Sub( q1, p1 ) );
	r.closest2 = b2MulAdd( p2, r.fraction2, b2Sub( q2, p2 ) );
	return r;
}

// src/manifold.c:36-70
B2_API b2Manifold b2CollideCircles( const b2Circle* circleA, b2Transform xfA, const b2Circle* circleB, b2Transform xfB )
{
	Manifold m;
	mf_clear( m );
	collide_circles( m, circleA->center, circleA->radius, xfA, circleB->center, circleB->radius, xfB );
	return to_public( m );
}

// src/manifold.c:73-135
B2_API b2Manifold b2CollideCapsuleAndCircle( const b2Capsule* capsuleA, b2Transform xfA, const b2Circle* circleB,
											 b2Transform xfB )
{
	Manifold m;
	mf_clear( m );
	collide_capsule_circle( m, capsuleA->center1, capsuleA->center2, capsuleA->radius, xfA, circleB->center, circleB->radius,
							xfB );
	return to_public( m );
}

// src/manifold.c:702-706
B2_API b2Manifold b2CollideSegmentAndCircle( const b2Segment* segmentA, b2Transform xfA, const b2Circle* circleB,
											 b2Transform xfB )
{
	Manifold m;
	mf_clear( m );
	collide_capsule_circle( m, segmentA->point1, segmentA->point2, 0.0f, xfA, circleB->center, circleB->radius, xfB );
	return to_public( m );
}

// src/manifold.c:137-253
B2_API b2Manifold b2CollidePolygonAndCircle( const b2Polygon* polygonA, b2Transform xfA, const b2Circle* circleB,
											 b2Transform xfB )
{
	Manifold m;
	mf_clear( m );
	Poly a = to_poly( polygonA );
	collide_polygon_circle( m, a, xfA, circleB->center, circleB->radius, xfB );
	return to_public( m );
}

// src/manifold.c:255-260
B2_API b2Manifold b2CollideCapsules( const b2Capsule* capsuleA, b2Transform xfA, const b2Capsule* capsuleB, b2Transform xfB )
{
	return polygons( capsule_poly( capsuleA->center1, capsuleA->center2, capsuleA->radius ), xfA,
					 capsule_poly( capsuleB->center1, capsuleB->center2, capsuleB->radius ), xfB );
}

// src/manifold.c:262-267
B2_API b2Manifold b2CollideSegmentAndCapsule( const b2Segment* segmentA, b2Transform xfA, const b2Capsule* capsuleB,
											  b2Transform xfB )
{
	return polygons( capsule_poly( segmentA->point1, segmentA->point2, 0.0f ), xfA,
					 capsule_poly( capsuleB->center1, capsuleB->center2, capsuleB->radius ), xfB );
}

// src/manifold.c:269-273
B2_API b2Manifold b2CollidePolygonAndCapsule( const b2Polygon* polygonA, b2Transform xfA, const b2Capsule* capsuleB,
											  b2Transform xfB )
{
	return polygons( to_poly( polygonA ), xfA, capsule_poly( capsuleB->center1, capsuleB->center2, capsuleB->radius ), xfB );
}

// src/manifold.c:470-700
B2_API b2Manifold b2CollidePolygons( const b2Polygon* polyA, b2Transform xfA, const b2Polygon* polyB, b2Transform xfB )
{
	return polygons( to_poly( polyA ), xfA, to_poly( polyB ), xfB );
}

// src/manifold.c:708-712
B2_API b2Manifold b2CollideSegmentAndPolygon( const b2Segment* segmentA, b2Transform xfA, const b2Polygon* polygonB,
											  b2Transform xfB )
{
	return polygons( capsule_poly( segmentA->point1, segmentA->point2, 0.0f ), xfA, to_poly( polygonB ), xfB );
}

// src/distance.c:109-120
B2_API b2DistanceProxy b2MakeProxy( const b2Vec2* vertices, int32_t count, float radius )
{
	return make_proxy( vertices, count, radius );
}

// src/distance.c:13-30
B2_API b2Transform b2GetSweepTransform( const b2Sweep* sweep, float time )
{
	return sweep_transform( *sweep, time );
}

// src/distance.c:459-623
B2_API b2DistanceOutput b2ShapeDistance( b2DistanceCache* cache, const b2DistanceInput* input, b2Simplex* simplexes, int simplexCapacity )
{
	return shape_distance( *cache, *input, simplexes, simplexCapacity );
}

// src/distance.c:999-1237
B2_API b2TOIOutput b2TimeOfImpact( const b2TOIInput* input )
{
	return time_of_impact( *input );
}

B2_API void b2mi_Atan2( const float* y, const float* x, float* out, int n )
{
	for ( int i = 0; i < n; ++i )
		out[i] = b2mi::fdlibm_atan2f( y[i], x[i] );
}

} // extern "C"
