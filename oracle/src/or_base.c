// CPU ORACLE — test infrastructure only (see or_internal.h).
// Containers, math helpers and geometry of the public API.
#include "or_internal.h"

#include <float.h>
#include <math.h>
#include <stdio.h>

// ---- assert hook (src/core.c:23-37) ----
static int or_default_assert( const char* condition, const char* fileName, int lineNumber )
{
	fprintf( stderr, "oracle assertion failed: %s, %s line %d\n", condition, fileName, lineNumber );
	return 1;
}

static b2AssertFcn* or_assert_fcn = or_default_assert;

B2_API void b2SetAssertFcn( b2AssertFcn* assertFcn )
{
	or_assert_fcn = assertFcn ? assertFcn : or_default_assert;
}

void or_assert_fail( const char* what, const char* file, int line )
{
	if ( or_assert_fcn( what, file, line ) )
		abort();
}

B2_API b2Version b2GetVersion( void )
{
	return ( b2Version ){ 3, 0, 0 };
}

// ---- id pool (src/id_pool.c:21-47) ----
int or_alloc_id( orIdPool* pool )
{
	if ( pool->freeArray.count > 0 )
		return pool->freeArray.data[--pool->freeArray.count];
	return pool->nextIndex++;
}

void or_free_id( orIdPool* pool, int id )
{
	OR_ASSERT( 0 <= id && id < pool->nextIndex );
	OR_PUSH( pool->freeArray, id );
}

// ---- bitset (src/bitset.c) ----
void or_bitset_set_count_and_clear( orBitSet* s, uint32_t bitCount )
{
	uint32_t blockCount = ( bitCount + 63 ) / 64;
	if ( s->blockCapacity < blockCount )
	{
		free( s->bits );
		uint32_t cap = blockCount + blockCount / 2 + 1;
		s->bits = calloc( cap, sizeof( uint64_t ) );
		s->blockCapacity = cap;
	}
	s->blockCount = blockCount;
	memset( s->bits, 0, blockCount * sizeof( uint64_t ) );
}

void or_bitset_grow( orBitSet* s, uint32_t blockCount )
{
	if ( blockCount > s->blockCapacity )
	{
		uint32_t cap = blockCount + blockCount / 2;
		uint64_t* bits = calloc( cap, sizeof( uint64_t ) );
		if ( s->bits )
			memcpy( bits, s->bits, s->blockCount * sizeof( uint64_t ) );
		free( s->bits );
		s->bits = bits;
		s->blockCapacity = cap;
	}
	for ( uint32_t b = s->blockCount; b < blockCount; ++b )
		s->bits[b] = 0;
	s->blockCount = blockCount;
}

void or_bitset_free( orBitSet* s )
{
	free( s->bits );
	memset( s, 0, sizeof( *s ) );
}

// ---- key set: open addressing, linear probing, backward-shift delete ----
static uint32_t or_hash64( uint64_t k )
{
	k ^= k >> 33;
	k *= 0xff51afd7ed558ccdULL;
	k ^= k >> 33;
	k *= 0xc4ceb9fe1a85ec53ULL;
	k ^= k >> 33;
	return (uint32_t)k;
}

// inserts an already-offset stored value (never 0) into a table with a free slot
static void or_keyset_insert_stored( orKeySet* s, uint64_t stored )
{
	uint32_t mask = s->capacity - 1, i = or_hash64( stored ) & mask;
	while ( s->keys[i] )
		i = ( i + 1 ) & mask;
	s->keys[i] = stored;
	s->count += 1;
}

static void or_keyset_rehash( orKeySet* s, uint32_t capacity )
{
	uint64_t* old = s->keys;
	uint32_t oldCap = s->capacity;
	s->keys = calloc( capacity, sizeof( uint64_t ) );
	s->capacity = capacity;
	s->count = 0;
	for ( uint32_t i = 0; i < oldCap; ++i )
		if ( old[i] )
			or_keyset_insert_stored( s, old[i] );
	free( old );
}

// keys are stored +1 so that key 0 is representable
bool or_keyset_add( orKeySet* s, uint64_t key )
{
	uint64_t k = key + 1;
	if ( 2 * ( s->count + 1 ) > s->capacity )
		or_keyset_rehash( s, s->capacity ? 2 * s->capacity : 64 );
	uint32_t mask = s->capacity - 1, i = or_hash64( k ) & mask;
	while ( s->keys[i] )
	{
		if ( s->keys[i] == k )
			return true;
		i = ( i + 1 ) & mask;
	}
	s->keys[i] = k;
	s->count += 1;
	return false;
}

bool or_keyset_contains( const orKeySet* s, uint64_t key )
{
	if ( s->capacity == 0 )
		return false;
	uint64_t k = key + 1;
	uint32_t mask = s->capacity - 1, i = or_hash64( k ) & mask;
	while ( s->keys[i] )
	{
		if ( s->keys[i] == k )
			return true;
		i = ( i + 1 ) & mask;
	}
	return false;
}

bool or_keyset_remove( orKeySet* s, uint64_t key )
{
	if ( s->capacity == 0 )
		return false;
	uint64_t k = key + 1;
	uint32_t mask = s->capacity - 1, i = or_hash64( k ) & mask;
	while ( s->keys[i] && s->keys[i] != k )
		i = ( i + 1 ) & mask;
	if ( s->keys[i] == 0 )
		return false;
	// backward-shift deletion
	uint32_t j = i;
	for ( ;; )
	{
		j = ( j + 1 ) & mask;
		if ( s->keys[j] == 0 )
			break;
		uint32_t home = or_hash64( s->keys[j] ) & mask;
		bool between = ( i <= j ) ? ( i < home && home <= j ) : ( i < home || home <= j );
		if ( !between )
		{
			s->keys[i] = s->keys[j];
			i = j;
		}
	}
	s->keys[i] = 0;
	s->count -= 1;
	return true;
}

void or_keyset_clear( orKeySet* s )
{
	if ( s->keys )
		memset( s->keys, 0, s->capacity * sizeof( uint64_t ) );
	s->count = 0;
}

void or_keyset_free( orKeySet* s )
{
	free( s->keys );
	memset( s, 0, sizeof( *s ) );
}

// ---- validity and normalization (src/math_functions.c:10-93) ----
B2_API bool b2IsValid( float a )
{
	if ( isnan( a ) )
		return false;
	if ( isinf( a ) )
		return false;
	return true;
}

B2_API bool b2Vec2_IsValid( b2Vec2 v )
{
	return b2IsValid( v.x ) && b2IsValid( v.y );
}

B2_API bool b2Rot_IsValid( b2Rot q )
{
	if ( b2IsValid( q.s ) == false || b2IsValid( q.c ) == false )
		return false;
	return b2IsNormalized( q );
}

B2_API bool b2AABB_IsValid( b2AABB a )
{
	b2Vec2 d = b2Sub( a.upperBound, a.lowerBound );
	bool valid = d.x >= 0.0f && d.y >= 0.0f;
	return valid && b2Vec2_IsValid( a.lowerBound ) && b2Vec2_IsValid( a.upperBound );
}

B2_API b2Vec2 b2Normalize( b2Vec2 v )
{
	float length = b2Length( v );
	if ( length < FLT_EPSILON )
		return b2Vec2_zero;
	float invLength = 1.0f / length;
	return ( b2Vec2 ){ invLength * v.x, invLength * v.y };
}

B2_API b2Vec2 b2GetLengthAndNormalize( float* length, b2Vec2 v )
{
	*length = b2Length( v );
	if ( *length < FLT_EPSILON )
		return b2Vec2_zero;
	float invLength = 1.0f / *length;
	return ( b2Vec2 ){ invLength * v.x, invLength * v.y };
}

// ---- geometry (src/geometry.c:127-429) ----
B2_API b2Polygon b2MakeBox( float hx, float hy )
{
	b2Polygon p = { 0 };
	p.count = 4;
	p.vertices[0] = ( b2Vec2 ){ -hx, -hy };
	p.vertices[1] = ( b2Vec2 ){ hx, -hy };
	p.vertices[2] = ( b2Vec2 ){ hx, hy };
	p.vertices[3] = ( b2Vec2 ){ -hx, hy };
	p.normals[0] = ( b2Vec2 ){ 0.0f, -1.0f };
	p.normals[1] = ( b2Vec2 ){ 1.0f, 0.0f };
	p.normals[2] = ( b2Vec2 ){ 0.0f, 1.0f };
	p.normals[3] = ( b2Vec2 ){ -1.0f, 0.0f };
	p.radius = 0.0f;
	p.centroid = b2Vec2_zero;
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
	b2Polygon p = { 0 };
	p.count = 4;
	p.vertices[0] = b2TransformPoint( xf, ( b2Vec2 ){ -hx, -hy } );
	p.vertices[1] = b2TransformPoint( xf, ( b2Vec2 ){ hx, -hy } );
	p.vertices[2] = b2TransformPoint( xf, ( b2Vec2 ){ hx, hy } );
	p.vertices[3] = b2TransformPoint( xf, ( b2Vec2 ){ -hx, hy } );
	p.normals[0] = b2RotateVector( xf.q, ( b2Vec2 ){ 0.0f, -1.0f } );
	p.normals[1] = b2RotateVector( xf.q, ( b2Vec2 ){ 1.0f, 0.0f } );
	p.normals[2] = b2RotateVector( xf.q, ( b2Vec2 ){ 0.0f, 1.0f } );
	p.normals[3] = b2RotateVector( xf.q, ( b2Vec2 ){ -1.0f, 0.0f } );
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

	b2Vec2 vertices[b2_maxPolygonVertices] = { 0 };
	int count = shape->count;
	float radius = shape->radius;
	if ( radius > 0.0f )
	{
		// rounded polygons: push vertices out along the averaged normal (geometry.c:302-313)
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
	return ( b2AABB ){ { p.x - r, p.y - r }, { p.x + r, p.y + r } };
}

B2_API b2AABB b2ComputeCapsuleAABB( const b2Capsule* shape, b2Transform xf )
{
	b2Vec2 v1 = b2TransformPoint( xf, shape->center1 );
	b2Vec2 v2 = b2TransformPoint( xf, shape->center2 );
	b2Vec2 r = { shape->radius, shape->radius };
	return ( b2AABB ){ b2Sub( b2Min( v1, v2 ), r ), b2Add( b2Max( v1, v2 ), r ) };
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
	return ( b2AABB ){ b2Sub( lower, r ), b2Add( upper, r ) };
}

B2_API b2AABB b2ComputeSegmentAABB( const b2Segment* shape, b2Transform xf )
{
	b2Vec2 v1 = b2TransformPoint( xf, shape->point1 );
	b2Vec2 v2 = b2TransformPoint( xf, shape->point2 );
	return ( b2AABB ){ b2Min( v1, v2 ), b2Max( v1, v2 ) };
}

// src/shape.c:457-478
b2AABB or_compute_shape_aabb( const orShape* shape, b2Transform xf )
{
	switch ( shape->type )
	{
		case b2_capsuleShape:
			return b2ComputeCapsuleAABB( &shape->capsule, xf );
		case b2_circleShape:
			return b2ComputeCircleAABB( &shape->circle, xf );
		case b2_polygonShape:
			return b2ComputePolygonAABB( &shape->polygon, xf );
		case b2_segmentShape:
			return b2ComputeSegmentAABB( &shape->segment, xf );
		default:
			return ( b2AABB ){ xf.p, xf.p };
	}
}

// The C library's atan2f over n argument pairs (b2RelativeAngle's arc tangent,
// include/box2d/math_functions.h:387-394), for checking the kernels' restatement of it.
B2_API void oracle_Atan2( const float* y, const float* x, float* out, int n )
{
	for ( int i = 0; i < n; ++i )
		out[i] = atan2f( y[i], x[i] );
}
