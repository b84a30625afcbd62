// Host-side geometry helpers of the public API: primitive construction, mass and AABBs
// (reference src/geometry.c:127-429, src/math_functions.c:10-93). These run at body and
// shape creation only; the float expressions keep the reference's operation order because
// mass and inertia feed every later step.
#include "world_internal.hpp"

#include <cfloat>
#include <cmath>

extern "C" {

B2_API bool b2IsValid( float a )
{
	return !std::isnan( a ) && !std::isinf( a );
}

B2_API bool b2Vec2_IsValid( b2Vec2 v )
{
	return b2IsValid( v.x ) && b2IsValid( v.y );
}

B2_API bool b2Rot_IsValid( b2Rot q )
{
	return b2IsValid( q.s ) && b2IsValid( q.c ) && b2IsNormalized( q );
}

B2_API bool b2AABB_IsValid( b2AABB a )
{
	b2Vec2 d = b2Sub( a.upperBound, a.lowerBound );
	return d.x >= 0.0f && d.y >= 0.0f && b2Vec2_IsValid( a.lowerBound ) && b2Vec2_IsValid( a.upperBound );
}

B2_API b2Vec2 b2Normalize( b2Vec2 v )
{
	float length = b2Length( v );
	if ( length < FLT_EPSILON )
		return b2Vec2{ 0.0f, 0.0f };
	float inv = 1.0f / length;
	return b2Vec2{ inv * v.x, inv * v.y };
}

B2_API b2Vec2 b2GetLengthAndNormalize( float* length, b2Vec2 v )
{
	*length = b2Length( v );
	if ( *length < FLT_EPSILON )
		return b2Vec2{ 0.0f, 0.0f };
	float inv = 1.0f / *length;
	return b2Vec2{ inv * v.x, inv * v.y };
}

B2_API b2Polygon b2MakeBox( float hx, float hy )
{
	b2Polygon p{};
	p.count = 4;
	p.vertices[0] = { -hx, -hy };
	p.vertices[1] = { hx, -hy };
	p.vertices[2] = { hx, hy };
	p.vertices[3] = { -hx, hy };
	p.normals[0] = { 0.0f, -1.0f };
	p.normals[1] = { 1.0f, 0.0f };
	p.normals[2] = { 0.0f, 1.0f };
	p.normals[3] = { -1.0f, 0.0f };
	p.radius = 0.0f;
	p.centroid = { 0.0f, 0.0f };
	return p;
}

B2_API b2Polygon b2MakeSquare( float h )
{
	return b2MakeBox( h, h );
}

B2_API b2Polygon b2MakeRoundedBox( float hx, float hy, float radius )
{
	b2Polygon p = b2MakeBox( hx, hy );
	p.radius = radius;
	return p;
}

B2_API b2Polygon b2MakeOffsetBox( float hx, float hy, b2Vec2 center, float angle )
{
	b2Transform xf = { center, b2MakeRot( angle ) };
	b2Polygon p{};
	p.count = 4;
	p.vertices[0] = b2TransformPoint( xf, b2Vec2{ -hx, -hy } );
	p.vertices[1] = b2TransformPoint( xf, b2Vec2{ hx, -hy } );
	p.vertices[2] = b2TransformPoint( xf, b2Vec2{ hx, hy } );
	p.vertices[3] = b2TransformPoint( xf, b2Vec2{ -hx, hy } );
	p.normals[0] = b2RotateVector( xf.q, b2Vec2{ 0.0f, -1.0f } );
	p.normals[1] = b2RotateVector( xf.q, b2Vec2{ 1.0f, 0.0f } );
	p.normals[2] = b2RotateVector( xf.q, b2Vec2{ 0.0f, 1.0f } );
	p.normals[3] = b2RotateVector( xf.q, b2Vec2{ -1.0f, 0.0f } );
	p.radius = 0.0f;
	p.centroid = center;
	return p;
}

B2_API b2MassData b2ComputeCircleMass( const b2Circle* shape, float density )
{
	float rr = shape->radius * shape->radius;
	b2MassData m;
	m.mass = density * b2_pi * rr;
	m.center = shape->center;
	m.rotationalInertia = m.mass * ( 0.5f * rr + b2Dot( shape->center, shape->center ) );
	return m;
}

B2_API b2MassData b2ComputeCapsuleMass( const b2Capsule* shape, float density )
{
	float radius = shape->radius;
	float rr = radius * radius;
	b2Vec2 p1 = shape->center1, p2 = shape->center2;
	float length = b2Length( b2Sub( p2, p1 ) );
	float ll = length * length;
	float circleMass = density * ( b2_pi * radius * radius );
	float boxMass = density * ( 2.0f * radius * length );
	b2MassData m;
	m.mass = circleMass + boxMass;
	m.center.x = 0.5f * ( p1.x + p2.x );
	m.center.y = 0.5f * ( p1.y + p2.y );
	float lc = 4.0f * radius / ( 3.0f * b2_pi );
	float h = 0.5f * length;
	float circleInertia = circleMass * ( 0.5f * rr + h * h + 2.0f * h * lc );
	float boxInertia = boxMass * ( 4.0f * rr + ll ) / 12.0f;
	m.rotationalInertia = circleInertia + boxInertia;
	m.rotationalInertia += m.mass * b2Dot( m.center, m.center );
	return m;
}

// triangle-fan integration about the first vertex (src/geometry.c:251-374)
B2_API b2MassData b2ComputePolygonMass( const b2Polygon* shape, float density )
{
	if ( shape->count == 1 )
	{
		b2Circle c = { shape->vertices[0], shape->radius };
		return b2ComputeCircleMass( &c, density );
	}
	if ( shape->count == 2 )
	{
		b2Capsule c = { shape->vertices[0], shape->vertices[1], shape->radius };
		return b2ComputeCapsuleMass( &c, density );
	}
	b2Vec2 vertices[b2_maxPolygonVertices] = {};
	int count = shape->count;
	float radius = shape->radius;
	if ( radius > 0.0f )
	{
		float sqrt2 = 1.412f;
		for ( int i = 0; i < count; ++i )
		{
			int j = i == 0 ? count - 1 : i - 1;
			b2Vec2 mid = b2Normalize( b2Add( shape->normals[j], shape->normals[i] ) );
			vertices[i] = b2MulAdd( shape->vertices[i], sqrt2 * radius, mid );
		}
	}
	else
	{
		for ( int i = 0; i < count; ++i )
			vertices[i] = shape->vertices[i];
	}
	b2Vec2 center = { 0.0f, 0.0f };
	float area = 0.0f, rotationalInertia = 0.0f;
	b2Vec2 r = vertices[0];
	const float inv3 = 1.0f / 3.0f;
	for ( int i = 1; i < count - 1; ++i )
	{
		b2Vec2 e1 = b2Sub( vertices[i], r );
		b2Vec2 e2 = b2Sub( vertices[i + 1], r );
		float D = b2Cross( e1, e2 );
		float triangleArea = 0.5f * D;
		area += triangleArea;
		center = b2MulAdd( center, triangleArea * inv3, b2Add( e1, e2 ) );
		float ex1 = e1.x, ey1 = e1.y, ex2 = e2.x, ey2 = e2.y;
		float intx2 = ex1 * ex1 + ex2 * ex1 + ex2 * ex2;
		float inty2 = ey1 * ey1 + ey2 * ey1 + ey2 * ey2;
		rotationalInertia += ( 0.25f * inv3 * D ) * ( intx2 + inty2 );
	}
	b2MassData m;
	m.mass = density * area;
	float invArea = 1.0f / area;
	center.x *= invArea;
	center.y *= invArea;
	m.center = b2Add( r, center );
	m.rotationalInertia = density * rotationalInertia;
	m.rotationalInertia += m.mass * ( b2Dot( m.center, m.center ) - b2Dot( center, center ) );
	return m;
}

B2_API b2AABB b2ComputeCircleAABB( const b2Circle* shape, b2Transform xf )
{
	b2Vec2 p = b2TransformPoint( xf, shape->center );
	float r = shape->radius;
	return b2AABB{ { p.x - r, p.y - r }, { p.x + r, p.y + r } };
}

B2_API b2AABB b2ComputeCapsuleAABB( const b2Capsule* shape, b2Transform xf )
{
	b2Vec2 v1 = b2TransformPoint( xf, shape->center1 );
	b2Vec2 v2 = b2TransformPoint( xf, shape->center2 );
	b2Vec2 r = { shape->radius, shape->radius };
	return b2AABB{ b2Sub( b2Min( v1, v2 ), r ), b2Add( b2Max( v1, v2 ), r ) };
}

B2_API b2AABB b2ComputePolygonAABB( const b2Polygon* shape, b2Transform xf )
{
	b2Vec2 lower = b2TransformPoint( xf, shape->vertices[0] );
	b2Vec2 upper = lower;
	for ( int i = 1; i < shape->count; ++i )
	{
		b2Vec2 v = b2TransformPoint( xf, shape->vertices[i] );
		lower = b2Min( lower, v );
		upper = b2Max( upper, v );
	}
	b2Vec2 r = { shape->radius, shape->radius };
	return b2AABB{ b2Sub( lower, r ), b2Add( upper, r ) };
}

B2_API b2AABB b2ComputeSegmentAABB( const b2Segment* shape, b2Transform xf )
{
	b2Vec2 v1 = b2TransformPoint( xf, shape->point1 );
	b2Vec2 v2 = b2TransformPoint( xf, shape->point2 );
	return b2AABB{ b2Min( v1, v2 ), b2Max( v1, v2 ) };
}

} // extern "C"
