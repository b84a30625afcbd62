// Geometry and contact-manifold types of the public API.
//
// Layouts follow the reference's include/box2d/collision.h (b2Polygon 84-140, b2Manifold
// 474-520) field for field, since b2ContactData and b2Shape_GetPolygon hand them across the
// C ABI by value.
#pragma once

#include "base.h"
#include "math_functions.h"

#include <stdbool.h>

#define b2_maxPolygonVertices 8

typedef struct b2MassData
{
	float mass;
	b2Vec2 center;
	float rotationalInertia;
} b2MassData;

typedef struct b2Circle
{
	b2Vec2 center;
	float radius;
} b2Circle;

typedef struct b2Capsule
{
	b2Vec2 center1;
	b2Vec2 center2;
	float radius;
} b2Capsule;

// Convex polygon, counter-clockwise, with an optional rounding radius.
typedef struct b2Polygon
{
	b2Vec2 vertices[b2_maxPolygonVertices];
	b2Vec2 normals[b2_maxPolygonVertices];
	b2Vec2 centroid;
	float radius;
	int32_t count;
} b2Polygon;

typedef struct b2Segment
{
	b2Vec2 point1;
	b2Vec2 point2;
} b2Segment;

typedef struct b2SegmentDistanceResult
{
	b2Vec2 closest1;
	b2Vec2 closest2;
	float fraction1;
	float fraction2;
	float distanceSquared;
} b2SegmentDistanceResult;

// One contact point. anchorA/anchorB are relative to the body centers of mass once a
// contact has been updated by the step; id encodes the clipping features (i << 8 | j).
typedef struct b2ManifoldPoint
{
	b2Vec2 point;
	b2Vec2 anchorA;
	b2Vec2 anchorB;
	float separation;
	float normalImpulse;
	float tangentImpulse;
	float maxNormalImpulse;
	float normalVelocity;
	uint16_t id;
	bool persisted;
} b2ManifoldPoint;

typedef struct b2Manifold
{
	b2ManifoldPoint points[2];
	b2Vec2 normal;
	int32_t pointCount;
} b2Manifold;

// Geometry helpers (reference src/geometry.c:127-429).
B2_API b2Polygon b2MakeBox( float hx, float hy );
B2_API b2Polygon b2MakeSquare( float h );
B2_API b2Polygon b2MakeRoundedBox( float hx, float hy, float radius );
B2_API b2Polygon b2MakeOffsetBox( float hx, float hy, b2Vec2 center, float angle );
B2_API b2MassData b2ComputeCircleMass( const b2Circle* shape, float density );
B2_API b2MassData b2ComputeCapsuleMass( const b2Capsule* shape, float density );
B2_API b2MassData b2ComputePolygonMass( const b2Polygon* shape, float density );
B2_API b2AABB b2ComputeCircleAABB( const b2Circle* shape, b2Transform transform );
B2_API b2AABB b2ComputeCapsuleAABB( const b2Capsule* shape, b2Transform transform );
B2_API b2AABB b2ComputePolygonAABB( const b2Polygon* shape, b2Transform transform );
B2_API b2AABB b2ComputeSegmentAABB( const b2Segment* shape, b2Transform transform );

// Segment closest points (reference src/distance.c:32-105).
B2_API b2SegmentDistanceResult b2SegmentDistance( b2Vec2 p1, b2Vec2 q1, b2Vec2 p2, b2Vec2 q2 );

// GJK distance and time of impact (reference include/box2d/collision.h:316-455,
// src/distance.c:13-30, 105-623, 785-1237). The types keep the reference's layouts: the
// continuous stage of the step builds them for every candidate pair.
typedef struct b2DistanceProxy
{
	b2Vec2 points[b2_maxPolygonVertices];
	int32_t count;
	float radius;
} b2DistanceProxy;

typedef struct b2DistanceCache
{
	uint16_t count;
	uint8_t indexA[3];
	uint8_t indexB[3];
} b2DistanceCache;

typedef struct b2DistanceInput
{
	b2DistanceProxy proxyA;
	b2DistanceProxy proxyB;
	b2Transform transformA;
	b2Transform transformB;
	bool useRadii;
} b2DistanceInput;

typedef struct b2DistanceOutput
{
	b2Vec2 pointA;
	b2Vec2 pointB;
	float distance;
	int32_t iterations;
	int32_t simplexCount;
} b2DistanceOutput;

typedef struct b2SimplexVertex
{
	b2Vec2 wA;
	b2Vec2 wB;
	b2Vec2 w;
	float a;
	int32_t indexA;
	int32_t indexB;
} b2SimplexVertex;

typedef struct b2Simplex
{
	b2SimplexVertex v1, v2, v3;
	int32_t count;
} b2Simplex;

typedef struct b2Sweep
{
	b2Vec2 localCenter;
	b2Vec2 c1;
	b2Vec2 c2;
	b2Rot q1;
	b2Rot q2;
} b2Sweep;

typedef struct b2TOIInput
{
	b2DistanceProxy proxyA;
	b2DistanceProxy proxyB;
	b2Sweep sweepA;
	b2Sweep sweepB;
	float tMax;
} b2TOIInput;

typedef enum b2TOIState
{
	b2_toiStateUnknown,
	b2_toiStateFailed,
	b2_toiStateOverlapped,
	b2_toiStateHit,
	b2_toiStateSeparated
} b2TOIState;

typedef struct b2TOIOutput
{
	b2TOIState state;
	float t;
} b2TOIOutput;

// reference include/box2d/collision.h:389, 413, 428, 455
B2_API b2DistanceOutput b2ShapeDistance( b2DistanceCache* cache, const b2DistanceInput* input, b2Simplex* simplexes,
										 int simplexCapacity );
B2_API b2DistanceProxy b2MakeProxy( const b2Vec2* vertices, int32_t count, float radius );
B2_API b2Transform b2GetSweepTransform( const b2Sweep* sweep, float time );
B2_API b2TOIOutput b2TimeOfImpact( const b2TOIInput* input );

// Narrowphase manifolds (reference src/manifold.c:14-712).
B2_API b2Manifold b2CollideCircles( const b2Circle* circleA, b2Transform xfA, const b2Circle* circleB, b2Transform xfB );
B2_API b2Manifold b2CollideCapsuleAndCircle( const b2Capsule* capsuleA, b2Transform xfA, const b2Circle* circleB,
											 b2Transform xfB );
B2_API b2Manifold b2CollideSegmentAndCircle( const b2Segment* segmentA, b2Transform xfA, const b2Circle* circleB,
											 b2Transform xfB );
B2_API b2Manifold b2CollidePolygonAndCircle( const b2Polygon* polygonA, b2Transform xfA, const b2Circle* circleB,
											 b2Transform xfB );
B2_API b2Manifold b2CollideCapsules( const b2Capsule* capsuleA, b2Transform xfA, const b2Capsule* capsuleB, b2Transform xfB );
B2_API b2Manifold b2CollideSegmentAndCapsule( const b2Segment* segmentA, b2Transform xfA, const b2Capsule* capsuleB,
											  b2Transform xfB );
B2_API b2Manifold b2CollidePolygonAndCapsule( const b2Polygon* polygonA, b2Transform xfA, const b2Capsule* capsuleB,
											  b2Transform xfB );
B2_API b2Manifold b2CollidePolygons( const b2Polygon* polyA, b2Transform xfA, const b2Polygon* polyB, b2Transform xfB );
B2_API b2Manifold b2CollideSegmentAndPolygon( const b2Segment* segmentA, b2Transform xfA, const b2Polygon* polygonB,
											  b2Transform xfB );

#define b2_defaultCategoryBits ( 0x00000001 )
#define b2_defaultMaskBits ( 0xFFFFFFFF )
