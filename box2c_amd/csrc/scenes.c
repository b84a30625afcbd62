// Benchmark and test scenes, written against the public C API only (include/box2d/box2d.h),
// so the same file is linked against this library, against the CPU oracle and against the
// reference build; each scene is then bit-for-bit the same input for all three.
//
// Scene definitions:
//  - large pyramid: the reference's benchmark/large_pyramid.c:7-55 layout, base parameterized.
//  - pyramid grid: rows x cols pyramids (base 10, box half-extent 0.5) spaced 12 m apart,
//    after benchmark/many_pyramids.c:7-30; every row stands on its own static box of one
//    ground body; each box x is jittered by U(-0.01, 0.01) from a PCG32 stream.
//  - tumbler: a kinematic square container spinning at 0.25 rad/s about its center (after
//    samples/sample_benchmark.cpp:407-467) filled with half boxes / half circles of size
//    0.125 on a seeded jittered 0.4 m grid.
//  - batch pyramid: one world of a batch; base drawn from {4..40}, friction U(0.4, 0.8).
//  - drop boxes: one world of the differentiable batch; boxes with half-extents U(0.2, 0.6)
//    dropped from heights U(1, 4) with N(0, 2) linear and N(0, 1) angular velocity.
//  - shape zoo: every supported shape pair (circle, capsule, polygon, rounded polygon,
//    segment) resting and colliding, for narrowphase coverage.
#include "box2d/box2d.h"

#include <math.h>
#include <string.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>

#if defined( __GNUC__ )
	#define SCENE_API __attribute__( ( visibility( "default" ) ) )
#else
	#define SCENE_API
#endif

// ---- PCG32 (O'Neill 2014, pcg32_random_r), seeded per (seed, stream) ----
typedef struct Pcg32
{
	uint64_t state;
	uint64_t inc;
} Pcg32;

static uint32_t pcg32_next( Pcg32* r )
{
	uint64_t old = r->state;
	r->state = old * 6364136223846793005ULL + r->inc;
	uint32_t xorshifted = (uint32_t)( ( ( old >> 18u ) ^ old ) >> 27u );
	uint32_t rot = (uint32_t)( old >> 59u );
	return ( xorshifted >> rot ) | ( xorshifted << ( ( -rot ) & 31u ) );
}

static Pcg32 pcg32_seed( uint64_t seed, uint64_t stream )
{
	Pcg32 r = { 0u, ( stream << 1u ) | 1u };
	pcg32_next( &r );
	r.state += seed;
	pcg32_next( &r );
	return r;
}

// uniform in [0, 1) with 24 random bits
static float pcg32_unit( Pcg32* r )
{
	return (float)( pcg32_next( r ) >> 8 ) * ( 1.0f / 16777216.0f );
}

static float pcg32_uniform( Pcg32* r, float lo, float hi )
{
	return lo + ( hi - lo ) * pcg32_unit( r );
}

// Box-Muller normal
static float pcg32_normal( Pcg32* r, float sigma )
{
	float u1 = pcg32_unit( r );
	float u2 = pcg32_unit( r );
	if ( u1 < 1.0e-7f )
		u1 = 1.0e-7f;
	return sigma * sqrtf( -2.0f * logf( u1 ) ) * cosf( 2.0f * b2_pi * u2 );
}

// ---- world factory (lets a batch API own the world) ----
// returns the world id packed as index1 | revision << 16 (plain integer for ctypes callers)
typedef uint32_t SceneWorldFactory( const b2WorldDef* def, void* context );
static SceneWorldFactory* s_factory = 0;
static void* s_factoryContext = 0;

SCENE_API void scene_set_world_factory( SceneWorldFactory* factory, void* context )
{
	s_factory = factory;
	s_factoryContext = context;
}

static b2WorldId scene_create_world( const b2WorldDef* def )
{
	if ( s_factory )
	{
		uint32_t packed = s_factory( def, s_factoryContext );
		b2WorldId id = { (uint16_t)( packed & 0xffffu ), (uint16_t)( packed >> 16 ) };
		return id;
	}
	return b2CreateWorld( def );
}

static void record( b2BodyId* out, int capacity, int index, b2BodyId id )
{
	if ( out && index < capacity )
		out[index] = id;
}

// ---- large pyramid (reference benchmark/large_pyramid.c) ----
SCENE_API int scene_large_pyramid( const b2WorldDef* def, int baseCount, b2WorldId* outWorld, b2BodyId* outBodies,
								   int capacity )
{
	b2WorldId worldId = scene_create_world( def );
	*outWorld = worldId;
	{
		b2BodyDef bodyDef = b2DefaultBodyDef();
		bodyDef.position = ( b2Vec2 ){ 0.0f, -1.0f };
		b2BodyId groundId = b2CreateBody( worldId, &bodyDef );
		b2Polygon box = b2MakeBox( 100.0f, 1.0f );
		b2ShapeDef shapeDef = b2DefaultShapeDef();
		b2CreatePolygonShape( groundId, &shapeDef, &box );
	}

	b2BodyDef bodyDef = b2DefaultBodyDef();
	bodyDef.type = b2_dynamicBody;
	bodyDef.enableSleep = false;
	b2ShapeDef shapeDef = b2DefaultShapeDef();
	shapeDef.density = 1.0f;

	float h = 0.5f;
	b2Polygon box = b2MakeSquare( h );
	float shift = 1.0f * h;
	int n = 0;
	for ( int i = 0; i < baseCount; ++i )
	{
		float y = ( 2.0f * i + 1.0f ) * shift;
		for ( int j = i; j < baseCount; ++j )
		{
			float x = ( i + 1.0f ) * shift + 2.0f * ( j - i ) * shift - h * baseCount;
			bodyDef.position = ( b2Vec2 ){ x, y };
			b2BodyId bodyId = b2CreateBody( worldId, &bodyDef );
			b2CreatePolygonShape( bodyId, &shapeDef, &box );
			record( outBodies, capacity, n++, bodyId );
		}
	}
	return n;
}

// ---- pyramid grid ----
SCENE_API int scene_pyramid_grid( const b2WorldDef* def, int rows, int cols, int baseCount, uint64_t seed,
								  b2WorldId* outWorld, b2BodyId* outBodies, int capacity )
{
	const float extent = 0.5f;
	const float pitch = 12.0f;
	Pcg32 rng = pcg32_seed( seed, 2 );

	b2WorldId worldId = scene_create_world( def );
	*outWorld = worldId;

	float halfWidth = 0.5f * pitch * cols + 2.0f;
	float x0 = -0.5f * pitch * ( cols - 1 );
	{
		b2BodyDef bodyDef = b2DefaultBodyDef();
		b2BodyId groundId = b2CreateBody( worldId, &bodyDef );
		b2ShapeDef shapeDef = b2DefaultShapeDef();
		for ( int r = 0; r < rows; ++r )
		{
			b2Polygon box = b2MakeOffsetBox( halfWidth, 0.5f, ( b2Vec2 ){ 0.0f, r * pitch - 0.5f }, 0.0f );
			b2CreatePolygonShape( groundId, &shapeDef, &box );
		}
	}

	b2BodyDef bodyDef = b2DefaultBodyDef();
	bodyDef.type = b2_dynamicBody;
	bodyDef.enableSleep = false;
	b2ShapeDef shapeDef = b2DefaultShapeDef();
	b2Polygon box = b2MakeSquare( extent );
	int n = 0;
	for ( int r = 0; r < rows; ++r )
	{
		float baseY = r * pitch;
		for ( int c = 0; c < cols; ++c )
		{
			float centerX = x0 + c * pitch;
			for ( int i = 0; i < baseCount; ++i )
			{
				float y = ( 2.0f * i + 1.0f ) * extent + baseY;
				for ( int j = i; j < baseCount; ++j )
				{
					float x = ( i + 1.0f ) * extent + 2.0f * ( j - i ) * extent - extent * baseCount + centerX;
					x += pcg32_uniform( &rng, -0.01f, 0.01f );
					bodyDef.position = ( b2Vec2 ){ x, y };
					b2BodyId bodyId = b2CreateBody( worldId, &bodyDef );
					b2CreatePolygonShape( bodyId, &shapeDef, &box );
					record( outBodies, capacity, n++, bodyId );
				}
			}
		}
	}
	return n;
}

// ---- tumbler ----
SCENE_API int scene_tumbler( const b2WorldDef* def, int shapeCount, uint64_t seed, b2WorldId* outWorld,
							 b2BodyId* outBodies, int capacity )
{
	const float pitch = 0.4f;
	const float size = 0.125f;
	Pcg32 rng = pcg32_seed( seed, 3 );

	int side = 1;
	while ( side * side < shapeCount )
		side += 1;
	float halfInterior = 0.5f * pitch * side + 0.5f;
	if ( halfInterior < 10.0f )
		halfInterior = 10.0f;

	b2WorldId worldId = scene_create_world( def );
	*outWorld = worldId;
	{
		b2BodyDef bodyDef = b2DefaultBodyDef();
		bodyDef.type = b2_kinematicBody;
		bodyDef.angularVelocity = 0.25f;
		bodyDef.enableSleep = false;
		b2BodyId containerId = b2CreateBody( worldId, &bodyDef );
		b2ShapeDef shapeDef = b2DefaultShapeDef();
		float t = 0.25f, H = halfInterior + t;
		b2Polygon wall;
		wall = b2MakeOffsetBox( t, H, ( b2Vec2 ){ H, 0.0f }, 0.0f );
		b2CreatePolygonShape( containerId, &shapeDef, &wall );
		wall = b2MakeOffsetBox( t, H, ( b2Vec2 ){ -H, 0.0f }, 0.0f );
		b2CreatePolygonShape( containerId, &shapeDef, &wall );
		wall = b2MakeOffsetBox( H, t, ( b2Vec2 ){ 0.0f, H }, 0.0f );
		b2CreatePolygonShape( containerId, &shapeDef, &wall );
		wall = b2MakeOffsetBox( H, t, ( b2Vec2 ){ 0.0f, -H }, 0.0f );
		b2CreatePolygonShape( containerId, &shapeDef, &wall );
	}

	b2BodyDef bodyDef = b2DefaultBodyDef();
	bodyDef.type = b2_dynamicBody;
	bodyDef.enableSleep = false;
	b2ShapeDef shapeDef = b2DefaultShapeDef();
	shapeDef.density = 1.0f;
	shapeDef.friction = 0.6f;
	b2Polygon box = b2MakeSquare( size );
	b2Circle circle = { { 0.0f, 0.0f }, size };
	float start = -0.5f * pitch * ( side - 1 );
	int n = 0;
	for ( int k = 0; k < shapeCount; ++k )
	{
		int ix = k % side, iy = k / side;
		float x = start + pitch * ix + pcg32_uniform( &rng, -0.05f, 0.05f );
		float y = start + pitch * iy + pcg32_uniform( &rng, -0.05f, 0.05f );
		bodyDef.position = ( b2Vec2 ){ x, y };
		b2BodyId bodyId = b2CreateBody( worldId, &bodyDef );
		if ( pcg32_next( &rng ) & 1u )
			b2CreatePolygonShape( bodyId, &shapeDef, &box );
		else
			b2CreateCircleShape( bodyId, &shapeDef, &circle );
		record( outBodies, capacity, n++, bodyId );
	}
	return n;
}

// ---- one world of the uneven pyramid batch ----
SCENE_API int scene_batch_pyramid( const b2WorldDef* def, uint64_t seed, int worldIndex, b2WorldId* outWorld,
								   b2BodyId* outBodies, int capacity, float* outFriction )
{
	Pcg32 rng = pcg32_seed( seed, 1000u + (uint64_t)worldIndex );
	int baseCount = 4 + (int)( pcg32_next( &rng ) % 37u );
	float friction = pcg32_uniform( &rng, 0.4f, 0.8f );
	if ( outFriction )
		*outFriction = friction;

	b2WorldId worldId = scene_create_world( def );
	*outWorld = worldId;
	{
		b2BodyDef bodyDef = b2DefaultBodyDef();
		bodyDef.position = ( b2Vec2 ){ 0.0f, -1.0f };
		b2BodyId groundId = b2CreateBody( worldId, &bodyDef );
		b2Polygon box = b2MakeBox( 0.5f * baseCount + 10.0f, 1.0f );
		b2ShapeDef shapeDef = b2DefaultShapeDef();
		shapeDef.friction = friction;
		b2CreatePolygonShape( groundId, &shapeDef, &box );
	}

	b2BodyDef bodyDef = b2DefaultBodyDef();
	bodyDef.type = b2_dynamicBody;
	bodyDef.enableSleep = false;
	b2ShapeDef shapeDef = b2DefaultShapeDef();
	shapeDef.friction = friction;
	float h = 0.5f;
	b2Polygon box = b2MakeSquare( h );
	int n = 0;
	for ( int i = 0; i < baseCount; ++i )
	{
		float y = ( 2.0f * i + 1.0f ) * h;
		for ( int j = i; j < baseCount; ++j )
		{
			float x = ( i + 1.0f ) * h + 2.0f * ( j - i ) * h - h * baseCount;
			bodyDef.position = ( b2Vec2 ){ x, y };
			b2BodyId bodyId = b2CreateBody( worldId, &bodyDef );
			b2CreatePolygonShape( bodyId, &shapeDef, &box );
			record( outBodies, capacity, n++, bodyId );
		}
	}
	return n;
}

// ---- one world of the differentiable drop batch ----
// outParams (optional) receives friction, restitution.
// Sliding/tumbling boxes dropped on a ground box. Per world: friction U(0.2, 1) on every shape
// (so every contact's mixed friction sqrt(f f) is f) and restitution U(0, 0.3) on the boxes (the
// ground has 0, so every contact's max-mixed restitution is the boxes'); per box half-extents
// U(0.2, 0.6), height U(1, 4), velocity N(0, 2) per axis, angular velocity N(0, 1).
// scene_drop_boxes_ex replaces the sampled friction / restitution when they are >= 0 and the
// sampled velocities when v0 / w0 are given (boxCount entries; v0 as x, y pairs).
SCENE_API int scene_drop_boxes_ex( const b2WorldDef* def, uint64_t seed, int worldIndex, int boxCount, float frictionOverride,
								   float restitutionOverride, const float* v0, const float* w0, b2WorldId* outWorld,
								   b2BodyId* outBodies, int capacity, float* outParams )
{
	Pcg32 rng = pcg32_seed( seed, 500000u + (uint64_t)worldIndex );
	float friction = pcg32_uniform( &rng, 0.2f, 1.0f );
	float restitution = pcg32_uniform( &rng, 0.0f, 0.3f );
	if ( frictionOverride >= 0.0f )
		friction = frictionOverride;
	if ( restitutionOverride >= 0.0f )
		restitution = restitutionOverride;
	if ( outParams )
	{
		outParams[0] = friction;
		outParams[1] = restitution;
	}

	b2WorldId worldId = scene_create_world( def );
	*outWorld = worldId;
	{
		b2BodyDef bodyDef = b2DefaultBodyDef();
		bodyDef.position = ( b2Vec2 ){ 0.0f, -1.0f };
		b2BodyId groundId = b2CreateBody( worldId, &bodyDef );
		b2Polygon box = b2MakeBox( 0.75f * boxCount + 10.0f, 1.0f );
		b2ShapeDef shapeDef = b2DefaultShapeDef();
		shapeDef.friction = friction;
		shapeDef.restitution = 0.0f;
		b2CreatePolygonShape( groundId, &shapeDef, &box );
	}

	int n = 0;
	for ( int i = 0; i < boxCount; ++i )
	{
		float hx = pcg32_uniform( &rng, 0.2f, 0.6f );
		float hy = pcg32_uniform( &rng, 0.2f, 0.6f );
		float y = pcg32_uniform( &rng, 1.0f, 4.0f );
		b2BodyDef bodyDef = b2DefaultBodyDef();
		bodyDef.type = b2_dynamicBody;
		bodyDef.enableSleep = false;
		bodyDef.position = ( b2Vec2 ){ ( i - 0.5f * ( boxCount - 1 ) ) * 1.5f, y };
		float vx = pcg32_normal( &rng, 2.0f ), vy = pcg32_normal( &rng, 2.0f ), w = pcg32_normal( &rng, 1.0f );
		bodyDef.linearVelocity = v0 ? ( b2Vec2 ){ v0[2 * i], v0[2 * i + 1] } : ( b2Vec2 ){ vx, vy };
		bodyDef.angularVelocity = w0 ? w0[i] : w;
		b2BodyId bodyId = b2CreateBody( worldId, &bodyDef );
		b2ShapeDef shapeDef = b2DefaultShapeDef();
		shapeDef.friction = friction;
		shapeDef.restitution = restitution;
		b2Polygon box = b2MakeBox( hx, hy );
		b2CreatePolygonShape( bodyId, &shapeDef, &box );
		record( outBodies, capacity, n++, bodyId );
	}
	return n;
}

SCENE_API int scene_drop_boxes( const b2WorldDef* def, uint64_t seed, int worldIndex, int boxCount, b2WorldId* outWorld,
								b2BodyId* outBodies, int capacity, float* outParams )
{
	return scene_drop_boxes_ex( def, seed, worldIndex, boxCount, -1.0f, -1.0f, NULL, NULL, outWorld, outBodies, capacity,
								outParams );
}

// ---- shape zoo: every supported shape pair ----
SCENE_API int scene_shape_zoo( const b2WorldDef* def, int copies, uint64_t seed, b2WorldId* outWorld, b2BodyId* outBodies,
							   int capacity )
{
	Pcg32 rng = pcg32_seed( seed, 7 );
	b2WorldId worldId = scene_create_world( def );
	*outWorld = worldId;

	// ground: a box, a slanted box and a segment
	{
		b2BodyDef bodyDef = b2DefaultBodyDef();
		b2BodyId groundId = b2CreateBody( worldId, &bodyDef );
		b2ShapeDef shapeDef = b2DefaultShapeDef();
		b2Polygon box = b2MakeOffsetBox( 40.0f, 0.5f, ( b2Vec2 ){ 0.0f, -0.5f }, 0.0f );
		b2CreatePolygonShape( groundId, &shapeDef, &box );
		b2Polygon ramp = b2MakeOffsetBox( 6.0f, 0.25f, ( b2Vec2 ){ -30.0f, 3.0f }, -0.3f );
		b2CreatePolygonShape( groundId, &shapeDef, &ramp );
		b2Segment seg = { { 20.0f, 2.0f }, { 38.0f, 4.0f } };
		b2CreateSegmentShape( groundId, &shapeDef, &seg );
	}

	b2ShapeDef shapeDef = b2DefaultShapeDef();
	int n = 0;
	for ( int k = 0; k < copies; ++k )
	{
		for ( int kind = 0; kind < 5; ++kind )
		{
			b2BodyDef bodyDef = b2DefaultBodyDef();
			bodyDef.type = b2_dynamicBody;
			bodyDef.enableSleep = false;
			float x = -36.0f + 3.6f * ( ( k * 5 + kind ) % 20 ) + pcg32_uniform( &rng, -0.3f, 0.3f );
			float y = 1.0f + 1.6f * ( ( k * 5 + kind ) / 20 ) + pcg32_uniform( &rng, 0.0f, 0.5f );
			bodyDef.position = ( b2Vec2 ){ x, y };
			bodyDef.rotation = b2MakeRot( pcg32_uniform( &rng, -1.0f, 1.0f ) );
			bodyDef.angularVelocity = pcg32_uniform( &rng, -2.0f, 2.0f );
			b2BodyId bodyId = b2CreateBody( worldId, &bodyDef );
			shapeDef.friction = pcg32_uniform( &rng, 0.2f, 0.9f );
			shapeDef.restitution = ( kind == 1 ) ? 0.3f : 0.0f;
			switch ( kind )
			{
				case 0:
				{
					b2Polygon box = b2MakeBox( pcg32_uniform( &rng, 0.2f, 0.7f ), pcg32_uniform( &rng, 0.2f, 0.7f ) );
					b2CreatePolygonShape( bodyId, &shapeDef, &box );
				}
				break;
				case 1:
				{
					b2Circle circle = { { 0.0f, 0.0f }, pcg32_uniform( &rng, 0.2f, 0.6f ) };
					b2CreateCircleShape( bodyId, &shapeDef, &circle );
				}
				break;
				case 2:
				{
					b2Capsule capsule = { { -0.4f, 0.0f }, { 0.4f, 0.0f }, pcg32_uniform( &rng, 0.15f, 0.35f ) };
					b2CreateCapsuleShape( bodyId, &shapeDef, &capsule );
				}
				break;
				case 3:
				{
					b2Polygon rbox = b2MakeRoundedBox( 0.3f, 0.2f, 0.1f );
					b2CreatePolygonShape( bodyId, &shapeDef, &rbox );
				}
				break;
				default:
				{
					// two shapes on one body: an offset box and a circle
					b2Polygon obox = b2MakeOffsetBox( 0.3f, 0.3f, ( b2Vec2 ){ 0.3f, 0.0f }, 0.2f );
					b2CreatePolygonShape( bodyId, &shapeDef, &obox );
					b2Circle circle = { { -0.3f, 0.1f }, 0.25f };
					b2CreateCircleShape( bodyId, &shapeDef, &circle );
				}
				break;
			}
			record( outBodies, capacity, n++, bodyId );
		}
	}
	return n;
}

// A regular polygon of `count` vertices (3..8) around the origin, built field by field (the
// vertices on a circle of `radius`, outward unit normals, the centroid at the origin).
static b2Polygon regular_polygon( int count, float radius, float roundness )
{
	b2Polygon p;
	memset( &p, 0, sizeof( p ) );
	p.count = count;
	for ( int i = 0; i < count; ++i )
	{
		float a = 6.2831853f * (float)i / (float)count;
		p.vertices[i] = ( b2Vec2 ){ radius * cosf( a ), radius * sinf( a ) };
	}
	for ( int i = 0; i < count; ++i )
	{
		b2Vec2 e = { p.vertices[( i + 1 ) % count].x - p.vertices[i].x, p.vertices[( i + 1 ) % count].y - p.vertices[i].y };
		float len = sqrtf( e.x * e.x + e.y * e.y );
		p.normals[i] = ( b2Vec2 ){ e.y / len, -e.x / len };
	}
	p.radius = roundness;
	return p;
}

// Polygons of 3 to 8 vertices (rounded ones too), boxes and circles piled on a ground box:
// narrowphase pairs beyond four vertex slots.
SCENE_API int scene_polygon_zoo( const b2WorldDef* def, int copies, uint64_t seed, b2WorldId* outWorld, b2BodyId* outBodies,
								 int capacity )
{
	Pcg32 rng = pcg32_seed( seed, 11 );
	b2WorldId worldId = scene_create_world( def );
	*outWorld = worldId;
	{
		b2BodyDef bodyDef = b2DefaultBodyDef();
		bodyDef.position = ( b2Vec2 ){ 0.0f, -1.0f };
		b2BodyId groundId = b2CreateBody( worldId, &bodyDef );
		b2ShapeDef shapeDef = b2DefaultShapeDef();
		b2Polygon box = b2MakeBox( 30.0f, 1.0f );
		b2CreatePolygonShape( groundId, &shapeDef, &box );
	}
	b2ShapeDef shapeDef = b2DefaultShapeDef();
	int n = 0;
	for ( int k = 0; k < copies; ++k )
	{
		for ( int kind = 0; kind < 8; ++kind )
		{
			b2BodyDef bodyDef = b2DefaultBodyDef();
			bodyDef.type = b2_dynamicBody;
			float x = -12.0f + 1.5f * ( ( k * 8 + kind ) % 16 ) + pcg32_uniform( &rng, -0.2f, 0.2f );
			float y = 0.8f + 1.4f * ( ( k * 8 + kind ) / 16 ) + pcg32_uniform( &rng, 0.0f, 0.4f );
			bodyDef.position = ( b2Vec2 ){ x, y };
			bodyDef.rotation = b2MakeRot( pcg32_uniform( &rng, -1.0f, 1.0f ) );
			b2BodyId bodyId = b2CreateBody( worldId, &bodyDef );
			shapeDef.friction = pcg32_uniform( &rng, 0.3f, 0.8f );
			float r = pcg32_uniform( &rng, 0.35f, 0.6f );
			if ( kind < 6 )
			{
				// 3, 5, 6, 7, 8 vertices; the sixth kind a rounded hexagon
				static const int counts[6] = { 3, 5, 6, 7, 8, 6 };
				b2Polygon poly = regular_polygon( counts[kind], r, kind == 5 ? 0.08f : 0.0f );
				b2CreatePolygonShape( bodyId, &shapeDef, &poly );
			}
			else if ( kind == 6 )
			{
				b2Polygon box = b2MakeBox( 0.6f * r, 0.4f * r );
				b2CreatePolygonShape( bodyId, &shapeDef, &box );
			}
			else
			{
				b2Circle circle = { { 0.0f, 0.0f }, 0.6f * r };
				b2CreateCircleShape( bodyId, &shapeDef, &circle );
			}
			record( outBodies, capacity, n++, bodyId );
		}
	}
	return n;
}

// ---- HelloWorld of the reference test suite (test/test_world.c:17-97) ----
SCENE_API int scene_hello( const b2WorldDef* def, b2WorldId* outWorld, b2BodyId* outBodies, int capacity )
{
	b2WorldId worldId = scene_create_world( def );
	*outWorld = worldId;
	b2BodyDef groundBodyDef = b2DefaultBodyDef();
	groundBodyDef.position = ( b2Vec2 ){ 0.0f, -10.0f };
	b2BodyId groundId = b2CreateBody( worldId, &groundBodyDef );
	b2Polygon groundBox = b2MakeBox( 50.0f, 10.0f );
	b2ShapeDef groundShapeDef = b2DefaultShapeDef();
	b2CreatePolygonShape( groundId, &groundShapeDef, &groundBox );

	b2BodyDef bodyDef = b2DefaultBodyDef();
	bodyDef.type = b2_dynamicBody;
	bodyDef.position = ( b2Vec2 ){ 0.0f, 4.0f };
	b2BodyId bodyId = b2CreateBody( worldId, &bodyDef );
	b2Polygon dynamicBox = b2MakeBox( 1.0f, 1.0f );
	b2ShapeDef shapeDef = b2DefaultShapeDef();
	shapeDef.density = 1.0f;
	shapeDef.friction = 0.3f;
	b2CreatePolygonShape( bodyId, &shapeDef, &dynamicBox );
	record( outBodies, capacity, 0, bodyId );
	return 1;
}

// The reference's determinism test scene (test/test_determinism.c:90-150): 10 columns of
// 10 rounded boxes (half-extent 0.45, radius 0.05, friction 0.3), each column tilted by a
// 0.2 m offset per row, on a 1000 x 1 ground box. The test steps it 100 times with 3
// substeps and sleep disabled.
SCENE_API int scene_tilted_stacks( const b2WorldDef* def, b2WorldId* outWorld, b2BodyId* outBodies, int capacity )
{
	b2WorldId worldId = scene_create_world( def );
	*outWorld = worldId;
	{
		b2BodyDef bd = b2DefaultBodyDef();
		bd.position = ( b2Vec2 ){ 0.0f, -1.0f };
		b2BodyId groundId = b2CreateBody( worldId, &bd );
		b2Polygon box = b2MakeBox( 1000.0f, 1.0f );
		b2ShapeDef sd = b2DefaultShapeDef();
		b2CreatePolygonShape( groundId, &sd, &box );
	}
	b2Polygon box = b2MakeRoundedBox( 0.45f, 0.45f, 0.05f );
	b2ShapeDef sd = b2DefaultShapeDef();
	sd.density = 1.0f;
	sd.friction = 0.3f;
	const int columns = 10, rows = 10;
	float offset = 0.2f, dx = 5.0f;
	float xroot = -0.5f * dx * ( columns - 1.0f );
	for ( int j = 0; j < columns; ++j )
	{
		float x = xroot + j * dx;
		for ( int i = 0; i < rows; ++i )
		{
			b2BodyDef bd = b2DefaultBodyDef();
			bd.type = b2_dynamicBody;
			bd.position = ( b2Vec2 ){ x + offset * i, 0.5f + 1.0f * i };
			b2BodyId bodyId = b2CreateBody( worldId, &bd );
			b2CreatePolygonShape( bodyId, &sd, &box );
			record( outBodies, capacity, j * rows + i, bodyId );
		}
	}
	return columns * rows;
}

// Overflow constraints: a dynamic plank on the ground carrying `count` small boxes, and a
// kinematic rake resting on the boxes. The plank and the rake each touch every box, and the
// graph colouring holds at most one of a body's contacts per colour, so the contacts beyond the
// twelfth go to the overflow set: those of the plank share a dynamic body, those of the rake a
// kinematic one. `rakeVx` is the rake's linear velocity x (-0.0 makes that component negative
// zero).
SCENE_API int scene_overflow( const b2WorldDef* def, int count, float rakeVx, b2WorldId* outWorld, b2BodyId* outBodies,
							  int capacity )
{
	b2WorldId worldId = scene_create_world( def );
	*outWorld = worldId;
	{
		b2BodyDef bd = b2DefaultBodyDef();
		bd.position = ( b2Vec2 ){ 0.0f, -1.0f };
		b2BodyId groundId = b2CreateBody( worldId, &bd );
		b2Polygon box = b2MakeBox( 100.0f, 1.0f );
		b2ShapeDef sd = b2DefaultShapeDef();
		b2CreatePolygonShape( groundId, &sd, &box );
	}
	float pitch = 0.7f, half = 0.3f;
	float length = 0.5f * pitch * count + 1.0f;
	b2ShapeDef sd = b2DefaultShapeDef();
	{
		b2BodyDef bd = b2DefaultBodyDef();
		bd.type = b2_dynamicBody;
		bd.position = ( b2Vec2 ){ 0.0f, 0.25f };
		b2BodyId plankId = b2CreateBody( worldId, &bd );
		b2Polygon plank = b2MakeBox( length, 0.25f );
		b2CreatePolygonShape( plankId, &sd, &plank );
		record( outBodies, capacity, 0, plankId );
	}
	b2Polygon box = b2MakeBox( half, half );
	float x0 = -0.5f * pitch * ( count - 1 );
	for ( int i = 0; i < count; ++i )
	{
		b2BodyDef bd = b2DefaultBodyDef();
		bd.type = b2_dynamicBody;
		bd.position = ( b2Vec2 ){ x0 + pitch * i, 0.5f + half };
		b2BodyId bodyId = b2CreateBody( worldId, &bd );
		b2CreatePolygonShape( bodyId, &sd, &box );
		record( outBodies, capacity, 1 + i, bodyId );
	}
	{
		b2BodyDef bd = b2DefaultBodyDef();
		bd.type = b2_kinematicBody;
		bd.position = ( b2Vec2 ){ 0.0f, 0.5f + 2.0f * half + 0.2f };
		bd.linearVelocity = ( b2Vec2 ){ rakeVx, 0.0f };
		b2BodyId rakeId = b2CreateBody( worldId, &bd );
		b2Polygon rake = b2MakeBox( length, 0.2f );
		b2CreatePolygonShape( rakeId, &sd, &rake );
	}
	return 1 + count;
}

// One dynamic body over (or on) a static ground box whose top face is the line y = 0
// (half-extents 50 x 1, centred at (0, -1)); used for closed-form checks of gravity,
// damping, friction and restitution.
//   shape   0 = circle of radius `size`, 3 = square box of half-extent `size`
//   ground  0 = no ground body
// The ground shape gets the same friction and restitution as the body, so the mixed
// values (sqrt(fA fB), max(rA, rB)) equal the body's.
SCENE_API int scene_single( const b2WorldDef* def, int shape, float size, float x, float y, float vx, float vy,
							float w, float friction, float restitution, float linearDamping, float angularDamping,
							int ground, b2WorldId* outWorld, b2BodyId* outBodies, int capacity )
{
	b2WorldId worldId = scene_create_world( def );
	*outWorld = worldId;
	if ( ground )
	{
		b2BodyDef bd = b2DefaultBodyDef();
		bd.position = ( b2Vec2 ){ 0.0f, -1.0f };
		b2BodyId groundId = b2CreateBody( worldId, &bd );
		b2Polygon box = b2MakeBox( 50.0f, 1.0f );
		b2ShapeDef sd = b2DefaultShapeDef();
		sd.friction = friction;
		sd.restitution = restitution;
		b2CreatePolygonShape( groundId, &sd, &box );
	}
	b2BodyDef bd = b2DefaultBodyDef();
	bd.type = b2_dynamicBody;
	bd.position = ( b2Vec2 ){ x, y };
	bd.linearVelocity = ( b2Vec2 ){ vx, vy };
	bd.angularVelocity = w;
	bd.linearDamping = linearDamping;
	bd.angularDamping = angularDamping;
	b2BodyId bodyId = b2CreateBody( worldId, &bd );
	b2ShapeDef sd = b2DefaultShapeDef();
	sd.density = 1.0f;
	sd.friction = friction;
	sd.restitution = restitution;
	if ( shape == 0 )
	{
		b2Circle c = { { 0.0f, 0.0f }, size };
		b2CreateCircleShape( bodyId, &sd, &c );
	}
	else
	{
		b2Polygon box = b2MakeBox( size, size );
		b2CreatePolygonShape( bodyId, &sd, &box );
	}
	record( outBodies, capacity, 0, bodyId );
	return 1;
}

// Steps a world n times (saves one ctypes crossing per step in tests and benches).
SCENE_API void scene_step_n( b2WorldId worldId, float timeStep, int subStepCount, int n )
{
	for ( int i = 0; i < n; ++i )
		b2World_Step( worldId, timeStep, subStepCount );
}

// ---- joint scenes ----

// The reference's own tumbler benchmark (benchmark/tumbler.c:7-81): a dynamic box container of
// density 50 on a revolute joint to a static ground body, its motor at 25 deg/s with a 1e8 N m
// torque limit, filled with a gridCount x gridCount grid of 0.125 m boxes at 0.4 m pitch (the
// reference's Release build uses 45).
SCENE_API int scene_joint_tumbler( const b2WorldDef* def, int gridCount, b2WorldId* outWorld, b2BodyId* outBodies, int capacity )
{
	b2WorldId worldId = scene_create_world( def );
	*outWorld = worldId;
	b2BodyDef groundDef = b2DefaultBodyDef();
	b2BodyId groundId = b2CreateBody( worldId, &groundDef );
	int n = 0;
	{
		b2BodyDef bodyDef = b2DefaultBodyDef();
		bodyDef.type = b2_dynamicBody;
		bodyDef.position = ( b2Vec2 ){ 0.0f, 10.0f };
		b2BodyId bodyId = b2CreateBody( worldId, &bodyDef );
		b2ShapeDef shapeDef = b2DefaultShapeDef();
		shapeDef.density = 50.0f;
		b2Polygon polygon = b2MakeOffsetBox( 0.5f, 10.0f, ( b2Vec2 ){ 10.0f, 0.0f }, 0.0f );
		b2CreatePolygonShape( bodyId, &shapeDef, &polygon );
		polygon = b2MakeOffsetBox( 0.5f, 10.0f, ( b2Vec2 ){ -10.0f, 0.0f }, 0.0f );
		b2CreatePolygonShape( bodyId, &shapeDef, &polygon );
		polygon = b2MakeOffsetBox( 10.0f, 0.5f, ( b2Vec2 ){ 0.0f, 10.0f }, 0.0f );
		b2CreatePolygonShape( bodyId, &shapeDef, &polygon );
		polygon = b2MakeOffsetBox( 10.0f, 0.5f, ( b2Vec2 ){ 0.0f, -10.0f }, 0.0f );
		b2CreatePolygonShape( bodyId, &shapeDef, &polygon );
		record( outBodies, capacity, n++, bodyId );

		b2RevoluteJointDef jd = b2DefaultRevoluteJointDef();
		jd.bodyIdA = groundId;
		jd.bodyIdB = bodyId;
		jd.localAnchorA = ( b2Vec2 ){ 0.0f, 10.0f };
		jd.localAnchorB = ( b2Vec2 ){ 0.0f, 0.0f };
		jd.referenceAngle = 0.0f;
		jd.motorSpeed = ( b2_pi / 180.0f ) * 25.0f;
		jd.maxMotorTorque = 1e8f;
		jd.enableMotor = true;
		b2CreateRevoluteJoint( worldId, &jd );
	}
	b2Polygon polygon = b2MakeBox( 0.125f, 0.125f );
	b2BodyDef bodyDef = b2DefaultBodyDef();
	bodyDef.type = b2_dynamicBody;
	b2ShapeDef shapeDef = b2DefaultShapeDef();
	float y = -0.2f * gridCount + 10.0f;
	for ( int i = 0; i < gridCount; ++i )
	{
		float x = -0.2f * gridCount;
		for ( int j = 0; j < gridCount; ++j )
		{
			bodyDef.position = ( b2Vec2 ){ x, y };
			b2BodyId bodyId = b2CreateBody( worldId, &bodyDef );
			b2CreatePolygonShape( bodyId, &shapeDef, &polygon );
			record( outBodies, capacity, n++, bodyId );
			x += 0.4f;
		}
		y += 0.4f;
	}
	return n;
}

// The reference's joint grid benchmark (benchmark/joint_grid.c:8-83): an N x N grid of 0.4 m
// circles, each joined to its upper and left neighbours by revolute joints, the middle seven of
// the top row static, circles not colliding with each other. `gravityY` replaces the benchmark's
// zero gravity when nonzero. Records the dynamic bodies.
SCENE_API int scene_joint_grid( const b2WorldDef* def, int N, float gravityY, b2WorldId* outWorld, b2BodyId* outBodies, int capacity )
{
	b2WorldDef wd = *def;
	wd.gravity = ( b2Vec2 ){ 0.0f, gravityY };
	b2WorldId worldId = scene_create_world( &wd );
	*outWorld = worldId;
	b2ShapeDef shapeDef = b2DefaultShapeDef();
	shapeDef.density = 1.0f;
	shapeDef.filter.categoryBits = 2;
	shapeDef.filter.maskBits = ~2u;
	b2Circle circle = { { 0.0f, 0.0f }, 0.4f };
	b2RevoluteJointDef jd = b2DefaultRevoluteJointDef();
	b2BodyDef bodyDef = b2DefaultBodyDef();
	bodyDef.enableSleep = false;
	b2BodyId* bodies = malloc( (size_t)N * N * sizeof( b2BodyId ) );
	int n = 0, index = 0;
	for ( int k = 0; k < N; ++k )
	{
		for ( int i = 0; i < N; ++i )
		{
			float fk = (float)k, fi = (float)i;
			if ( k >= N / 2 - 3 && k <= N / 2 + 3 && i == 0 )
				bodyDef.type = b2_staticBody;
			else
				bodyDef.type = b2_dynamicBody;
			bodyDef.position = ( b2Vec2 ){ fk, -fi };
			b2BodyId body = b2CreateBody( worldId, &bodyDef );
			b2CreateCircleShape( body, &shapeDef, &circle );
			if ( i > 0 )
			{
				jd.bodyIdA = bodies[index - 1];
				jd.bodyIdB = body;
				jd.localAnchorA = ( b2Vec2 ){ 0.0f, -0.5f };
				jd.localAnchorB = ( b2Vec2 ){ 0.0f, 0.5f };
				b2CreateRevoluteJoint( worldId, &jd );
			}
			if ( k > 0 )
			{
				jd.bodyIdA = bodies[index - N];
				jd.bodyIdB = body;
				jd.localAnchorA = ( b2Vec2 ){ 0.5f, 0.0f };
				jd.localAnchorB = ( b2Vec2 ){ -0.5f, 0.0f };
				b2CreateRevoluteJoint( worldId, &jd );
			}
			bodies[index++] = body;
			if ( bodyDef.type == b2_dynamicBody )
				record( outBodies, capacity, n++, body );
		}
	}
	free( bodies );
	return n;
}

static b2BodyId zoo_box( b2WorldId worldId, b2BodyType type, float x, float y, float hx, float hy, float angle, b2BodyId* outBodies,
						 int capacity, int* n )
{
	b2BodyDef bd = b2DefaultBodyDef();
	bd.type = type;
	bd.position = ( b2Vec2 ){ x, y };
	bd.rotation = ( b2Rot ){ cosf( angle ), sinf( angle ) };
	b2BodyId id = b2CreateBody( worldId, &bd );
	b2ShapeDef sd = b2DefaultShapeDef();
	b2Polygon box = b2MakeBox( hx, hy );
	b2CreatePolygonShape( id, &sd, &box );
	record( outBodies, capacity, ( *n )++, id );
	return id;
}

// Every joint type the step solves, with springs, limits and motors, joints to static and
// kinematic bodies, collideConnected on and off, and a hub body with `spokes` joints (beyond the
// eleven graph colours, so the rest are overflow joints), all falling or swinging into contact
// with boxes and the ground. `types` is a bitmask over b2JointType of the joint kinds to include.
SCENE_API int scene_joint_zoo( const b2WorldDef* def, int spokes, uint32_t types, uint64_t seed, b2WorldId* outWorld,
							   b2BodyId* outBodies, int capacity )
{
	b2WorldId worldId = scene_create_world( def );
	*outWorld = worldId;
	Pcg32 rng = pcg32_seed( seed, 11 );
	int n = 0;
	b2BodyId ground;
	{
		b2BodyDef bd = b2DefaultBodyDef();
		bd.position = ( b2Vec2 ){ 0.0f, -1.0f };
		ground = b2CreateBody( worldId, &bd );
		b2Polygon box = b2MakeBox( 60.0f, 1.0f );
		b2ShapeDef sd = b2DefaultShapeDef();
		b2CreatePolygonShape( ground, &sd, &box );
	}
	// a few boxes on the ground for the chains to hit
	for ( int i = 0; i < 6; ++i )
		zoo_box( worldId, b2_dynamicBody, -14.0f + 0.9f * i, 0.5f, 0.4f, 0.5f, 0.0f, outBodies, capacity, &n );

	if ( types & ( 1u << b2_revoluteJoint ) )
	{
		// a pendulum chain from a static anchor: a limited link, a sprung link, a motorized link
		b2BodyId prev = ground;
		b2Vec2 anchor = { -12.0f, 9.0f };
		for ( int i = 0; i < 8; ++i )
		{
			b2BodyId link = zoo_box( worldId, b2_dynamicBody, anchor.x + 0.5f + 1.0f * i, anchor.y, 0.5f, 0.125f,
									 pcg32_uniform( &rng, -0.05f, 0.05f ), outBodies, capacity, &n );
			b2RevoluteJointDef jd = b2DefaultRevoluteJointDef();
			jd.bodyIdA = prev;
			jd.bodyIdB = link;
			jd.localAnchorA = i == 0 ? ( b2Vec2 ){ anchor.x, anchor.y + 1.0f } : ( b2Vec2 ){ 0.5f, 0.0f };
			jd.localAnchorB = ( b2Vec2 ){ -0.5f, 0.0f };
			if ( i == 2 )
			{
				jd.enableLimit = true;
				jd.lowerAngle = -0.25f * b2_pi;
				jd.upperAngle = 0.125f * b2_pi;
			}
			if ( i == 4 )
			{
				jd.enableSpring = true;
				jd.hertz = 2.0f;
				jd.dampingRatio = 0.3f;
			}
			if ( i == 6 )
			{
				jd.enableMotor = true;
				jd.motorSpeed = 1.5f;
				jd.maxMotorTorque = 5.0f;
			}
			b2CreateRevoluteJoint( worldId, &jd );
			prev = link;
		}
		// a wheel on a static axle with a motor and a limit, and two overlapping boxes that
		// collide through their joint
		b2BodyId wheel = zoo_box( worldId, b2_dynamicBody, 4.0f, 3.0f, 0.8f, 0.8f, 0.3f, outBodies, capacity, &n );
		b2RevoluteJointDef wj = b2DefaultRevoluteJointDef();
		wj.bodyIdA = ground;
		wj.bodyIdB = wheel;
		wj.localAnchorA = ( b2Vec2 ){ 4.0f, 4.0f };
		wj.referenceAngle = 0.3f;
		wj.enableMotor = true;
		wj.motorSpeed = -2.0f;
		wj.maxMotorTorque = 50.0f;
		wj.enableLimit = true;
		wj.lowerAngle = -1.0f;
		wj.upperAngle = 1.0f;
		b2CreateRevoluteJoint( worldId, &wj );
		b2BodyId a = zoo_box( worldId, b2_dynamicBody, 7.0f, 1.0f, 0.5f, 0.5f, 0.0f, outBodies, capacity, &n );
		b2BodyId b = zoo_box( worldId, b2_dynamicBody, 7.6f, 1.2f, 0.5f, 0.5f, 0.2f, outBodies, capacity, &n );
		b2RevoluteJointDef cj = b2DefaultRevoluteJointDef();
		cj.bodyIdA = a;
		cj.bodyIdB = b;
		cj.localAnchorA = ( b2Vec2 ){ 0.3f, 0.0f };
		cj.localAnchorB = ( b2Vec2 ){ -0.3f, -0.2f };
		cj.collideConnected = true;
		b2CreateRevoluteJoint( worldId, &cj );
		// a box hanging from a kinematic carriage moving sideways
		b2BodyDef kd = b2DefaultBodyDef();
		kd.type = b2_kinematicBody;
		kd.position = ( b2Vec2 ){ 12.0f, 8.0f };
		kd.linearVelocity = ( b2Vec2 ){ -0.5f, 0.0f };
		kd.angularVelocity = 0.2f;
		b2BodyId carriage = b2CreateBody( worldId, &kd );
		b2ShapeDef sd = b2DefaultShapeDef();
		b2Polygon cbox = b2MakeBox( 1.0f, 0.2f );
		b2CreatePolygonShape( carriage, &sd, &cbox );
		b2BodyId hanging = zoo_box( worldId, b2_dynamicBody, 12.0f, 6.5f, 0.3f, 0.6f, 0.1f, outBodies, capacity, &n );
		b2RevoluteJointDef kj = b2DefaultRevoluteJointDef();
		kj.bodyIdA = carriage;
		kj.bodyIdB = hanging;
		kj.localAnchorA = ( b2Vec2 ){ 0.0f, -0.3f };
		kj.localAnchorB = ( b2Vec2 ){ 0.0f, 1.2f };
		b2CreateRevoluteJoint( worldId, &kj );
	}

	if ( types & ( 1u << b2_weldJoint ) )
	{
		// a welded beam (rigid welds and one soft weld) falling onto the ground boxes
		b2BodyId prev = b2_nullBodyId;
		for ( int i = 0; i < 6; ++i )
		{
			b2BodyId seg = zoo_box( worldId, b2_dynamicBody, -15.0f + 0.8f * i, 4.0f + 0.1f * i, 0.4f, 0.15f, 0.05f, outBodies,
									capacity, &n );
			if ( i > 0 )
			{
				b2WeldJointDef wd = b2DefaultWeldJointDef();
				wd.bodyIdA = prev;
				wd.bodyIdB = seg;
				wd.localAnchorA = ( b2Vec2 ){ 0.4f, 0.0f };
				wd.localAnchorB = ( b2Vec2 ){ -0.4f, 0.0f };
				wd.referenceAngle = 0.0f;
				if ( i == 3 )
				{
					wd.linearHertz = 5.0f;
					wd.linearDampingRatio = 0.7f;
					wd.angularHertz = 3.0f;
					wd.angularDampingRatio = 0.5f;
				}
				b2CreateWeldJoint( worldId, &wd );
			}
			prev = seg;
		}
		// a box welded to the ground at an angle
		b2BodyId tilted = zoo_box( worldId, b2_dynamicBody, 16.0f, 2.0f, 0.3f, 1.0f, 0.4f, outBodies, capacity, &n );
		b2WeldJointDef gd = b2DefaultWeldJointDef();
		gd.bodyIdA = ground;
		gd.bodyIdB = tilted;
		gd.localAnchorA = ( b2Vec2 ){ 16.0f, 2.0f };
		gd.localAnchorB = ( b2Vec2 ){ 0.0f, -1.0f };
		gd.referenceAngle = 0.4f;
		b2CreateWeldJoint( worldId, &gd );
	}

	if ( types & ( 1u << b2_distanceJoint ) )
	{
		// a rope of boxes from a static anchor: rigid links, a sprung and limited link, a motorized
		// sprung link
		b2BodyId prev = ground;
		for ( int i = 0; i < 5; ++i )
		{
			b2BodyId link = zoo_box( worldId, b2_dynamicBody, -20.0f + 1.1f * ( i + 1 ), 12.0f, 0.2f, 0.2f, 0.0f, outBodies, capacity, &n );
			b2DistanceJointDef dd = b2DefaultDistanceJointDef();
			dd.bodyIdA = prev;
			dd.bodyIdB = link;
			dd.localAnchorA = i == 0 ? ( b2Vec2 ){ -20.0f, 13.0f } : ( b2Vec2 ){ 0.2f, 0.0f };
			dd.localAnchorB = ( b2Vec2 ){ -0.2f, 0.0f };
			dd.length = 0.7f;
			if ( i == 2 )
			{
				dd.enableSpring = true;
				dd.hertz = 3.0f;
				dd.dampingRatio = 0.2f;
				dd.enableLimit = true;
				dd.minLength = 0.5f;
				dd.maxLength = 1.0f;
			}
			if ( i == 3 )
			{
				dd.enableSpring = true;
				dd.hertz = 0.0f;
				dd.enableMotor = true;
				dd.motorSpeed = 0.5f;
				dd.maxMotorForce = 20.0f;
			}
			b2CreateDistanceJoint( worldId, &dd );
			prev = link;
		}
	}

	if ( types & ( 1u << b2_prismaticJoint ) )
	{
		// a slider on a static rail with limits and a motor, and a sprung slider between two boxes
		b2BodyId slider = zoo_box( worldId, b2_dynamicBody, -6.0f, 2.0f, 0.5f, 0.3f, 0.0f, outBodies, capacity, &n );
		b2PrismaticJointDef pd = b2DefaultPrismaticJointDef();
		pd.bodyIdA = ground;
		pd.bodyIdB = slider;
		pd.localAnchorA = ( b2Vec2 ){ -6.0f, 3.0f };
		pd.localAxisA = ( b2Vec2 ){ 1.0f, 0.2f };
		pd.enableLimit = true;
		pd.lowerTranslation = -2.0f;
		pd.upperTranslation = 1.5f;
		pd.enableMotor = true;
		pd.motorSpeed = 1.0f;
		pd.maxMotorForce = 50.0f;
		b2CreatePrismaticJoint( worldId, &pd );
		b2BodyId base = zoo_box( worldId, b2_dynamicBody, -3.0f, 1.0f, 0.6f, 0.3f, 0.0f, outBodies, capacity, &n );
		b2BodyId top = zoo_box( worldId, b2_dynamicBody, -3.0f, 2.5f, 0.3f, 0.3f, 0.1f, outBodies, capacity, &n );
		b2PrismaticJointDef sd = b2DefaultPrismaticJointDef();
		sd.bodyIdA = base;
		sd.bodyIdB = top;
		sd.localAnchorA = ( b2Vec2 ){ 0.0f, 1.2f };
		sd.localAxisA = ( b2Vec2 ){ 0.0f, 1.0f };
		sd.referenceAngle = 0.1f;
		sd.enableSpring = true;
		sd.hertz = 2.0f;
		sd.dampingRatio = 0.5f;
		b2CreatePrismaticJoint( worldId, &sd );
	}

	if ( types & ( 1u << b2_wheelJoint ) )
	{
		// a car: a chassis on two wheels with suspension springs, a driven rear wheel and a limited
		// front wheel
		b2BodyId chassis = zoo_box( worldId, b2_dynamicBody, 22.0f, 2.0f, 1.2f, 0.25f, 0.0f, outBodies, capacity, &n );
		for ( int k = 0; k < 2; ++k )
		{
			b2BodyDef bd = b2DefaultBodyDef();
			bd.type = b2_dynamicBody;
			bd.position = ( b2Vec2 ){ 22.0f + ( k == 0 ? -0.9f : 0.9f ), 1.2f };
			b2BodyId wheel = b2CreateBody( worldId, &bd );
			b2ShapeDef sd = b2DefaultShapeDef();
			b2Circle c = { { 0.0f, 0.0f }, 0.4f };
			b2CreateCircleShape( wheel, &sd, &c );
			record( outBodies, capacity, n++, wheel );
			b2WheelJointDef wd = b2DefaultWheelJointDef();
			wd.bodyIdA = chassis;
			wd.bodyIdB = wheel;
			wd.localAnchorA = ( b2Vec2 ){ k == 0 ? -0.9f : 0.9f, -0.8f };
			wd.localAxisA = ( b2Vec2 ){ 0.0f, 1.0f };
			wd.hertz = 4.0f;
			wd.dampingRatio = 0.7f;
			if ( k == 0 )
			{
				wd.enableMotor = true;
				wd.motorSpeed = -5.0f;
				wd.maxMotorTorque = 20.0f;
			}
			else
			{
				wd.enableLimit = true;
				wd.lowerTranslation = -0.25f;
				wd.upperTranslation = 0.25f;
			}
			b2CreateWheelJoint( worldId, &wd );
		}
	}

	if ( types & ( 1u << b2_motorJoint ) )
	{
		// a box driven toward an offset and angle from the ground body
		b2BodyId driven = zoo_box( worldId, b2_dynamicBody, 26.0f, 1.0f, 0.4f, 0.4f, 0.0f, outBodies, capacity, &n );
		b2MotorJointDef md = b2DefaultMotorJointDef();
		md.bodyIdA = ground;
		md.bodyIdB = driven;
		md.linearOffset = ( b2Vec2 ){ 27.0f, 4.0f };
		md.angularOffset = 0.5f;
		md.maxForce = 500.0f;
		md.maxTorque = 100.0f;
		md.correctionFactor = 0.3f;
		b2CreateMotorJoint( worldId, &md );
	}

	if ( types & ( 1u << b2_mouseJoint ) )
	{
		// a box dragged toward a target above it
		b2BodyId dragged = zoo_box( worldId, b2_dynamicBody, 30.0f, 0.5f, 0.5f, 0.5f, 0.2f, outBodies, capacity, &n );
		b2MouseJointDef md = b2DefaultMouseJointDef();
		md.bodyIdA = ground;
		md.bodyIdB = dragged;
		md.target = ( b2Vec2 ){ 31.0f, 5.0f };
		md.hertz = 5.0f;
		md.dampingRatio = 0.7f;
		md.maxForce = 200.0f;
		b2CreateMouseJoint( worldId, &md );
	}

	if ( spokes > 0 )
	{
		// a hub joined to `spokes` bars around it: joints past the eleventh colour overflow
		b2BodyId hub = zoo_box( worldId, b2_dynamicBody, 0.0f, 10.0f, 0.6f, 0.6f, 0.0f, outBodies, capacity, &n );
		for ( int i = 0; i < spokes; ++i )
		{
			float a = 2.0f * b2_pi * i / spokes;
			b2Vec2 dir = { cosf( a ), sinf( a ) };
			b2BodyId bar = zoo_box( worldId, b2_dynamicBody, 1.6f * dir.x, 10.0f + 1.6f * dir.y, 0.5f, 0.08f, a, outBodies, capacity,
									&n );
			b2Vec2 pivot = { 0.9f * dir.x, 10.0f + 0.9f * dir.y };
			if ( ( types & ( 1u << b2_weldJoint ) ) && ( i % 3 ) == 2 )
			{
				b2WeldJointDef wd = b2DefaultWeldJointDef();
				wd.bodyIdA = hub;
				wd.bodyIdB = bar;
				wd.localAnchorA = ( b2Vec2 ){ pivot.x, pivot.y - 10.0f };
				wd.localAnchorB = ( b2Vec2 ){ -0.7f, 0.0f };
				wd.referenceAngle = a;
				wd.angularHertz = i % 2 ? 4.0f : 0.0f;
				wd.angularDampingRatio = 0.5f;
				b2CreateWeldJoint( worldId, &wd );
			}
			else
			{
				b2RevoluteJointDef jd = b2DefaultRevoluteJointDef();
				jd.bodyIdA = hub;
				jd.bodyIdB = bar;
				jd.localAnchorA = ( b2Vec2 ){ pivot.x, pivot.y - 10.0f };
				jd.localAnchorB = ( b2Vec2 ){ -0.7f, 0.0f };
				jd.referenceAngle = a;
				jd.enableLimit = ( i % 2 ) == 0;
				jd.lowerAngle = -0.3f;
				jd.upperAngle = 0.3f;
				jd.enableSpring = ( i % 4 ) == 1;
				jd.hertz = 1.5f;
				jd.dampingRatio = 0.2f;
				b2CreateRevoluteJoint( worldId, &jd );
			}
		}
	}
	return n;
}

// Continuous collision: `count` bodies of every shape thrown down at 40-160 m/s (and up to 30 m/s
// sideways, spinning) at thin static shapes of one ground body: a 0.1 m floor, a slanted plank, a
// small circle and two segment walls, all thin enough that a body crosses them within a step. A
// long bar spins at 40 rad/s on a revolute motor about a point of the ground body; its tips sweep
// the floor, which its joint keeps out of its time of impact. The first `bullets` thrown bodies
// are bullets.
SCENE_API int scene_continuous( const b2WorldDef* def, int count, uint64_t seed, int bullets, b2WorldId* outWorld, b2BodyId* outBodies,
								int capacity )
{
	b2WorldId worldId = scene_create_world( def );
	*outWorld = worldId;
	Pcg32 rng = pcg32_seed( seed, 13 );
	int n = 0;
	b2BodyDef gd = b2DefaultBodyDef();
	b2BodyId ground = b2CreateBody( worldId, &gd );
	b2ShapeDef sd = b2DefaultShapeDef();
	{
		b2Polygon floor = b2MakeOffsetBox( 30.0f, 0.05f, ( b2Vec2 ){ 0.0f, 0.0f }, 0.0f );
		b2CreatePolygonShape( ground, &sd, &floor );
		b2Polygon plank = b2MakeOffsetBox( 4.0f, 0.05f, ( b2Vec2 ){ 10.0f, 6.0f }, 0.5f );
		b2CreatePolygonShape( ground, &sd, &plank );
		b2Circle post = { { -10.0f, 4.0f }, 0.3f };
		b2CreateCircleShape( ground, &sd, &post );
		b2Segment left = { { -30.0f, 0.0f }, { -30.0f, 40.0f } };
		b2Segment right = { { 30.0f, 40.0f }, { 30.0f, 0.0f } };
		b2CreateSegmentShape( ground, &sd, &left );
		b2CreateSegmentShape( ground, &sd, &right );
	}
	{
		b2BodyDef bd = b2DefaultBodyDef();
		bd.type = b2_dynamicBody;
		bd.position = ( b2Vec2 ){ 0.0f, 3.05f };
		b2BodyId bar = b2CreateBody( worldId, &bd );
		b2Polygon box = b2MakeBox( 3.0f, 0.05f );
		b2CreatePolygonShape( bar, &sd, &box );
		b2RevoluteJointDef jd = b2DefaultRevoluteJointDef();
		jd.bodyIdA = ground;
		jd.bodyIdB = bar;
		jd.localAnchorA = ( b2Vec2 ){ 0.0f, 3.05f };
		jd.localAnchorB = ( b2Vec2 ){ 0.0f, 0.0f };
		jd.enableMotor = true;
		jd.motorSpeed = 40.0f;
		jd.maxMotorTorque = 1.0e6f;
		b2CreateRevoluteJoint( worldId, &jd );
		if ( n < capacity )
			outBodies[n++] = bar;
	}
	for ( int i = 0; i < count; ++i )
	{
		b2BodyDef bd = b2DefaultBodyDef();
		bd.type = b2_dynamicBody;
		bd.isBullet = i < bullets;
		bd.position = ( b2Vec2 ){ pcg32_uniform( &rng, -25.0f, 25.0f ), pcg32_uniform( &rng, 4.0f, 20.0f ) };
		bd.rotation = b2MakeRot( pcg32_uniform( &rng, -3.0f, 3.0f ) );
		bd.linearVelocity = ( b2Vec2 ){ pcg32_uniform( &rng, -30.0f, 30.0f ), pcg32_uniform( &rng, -160.0f, -40.0f ) };
		bd.angularVelocity = pcg32_uniform( &rng, -20.0f, 20.0f );
		b2BodyId body = b2CreateBody( worldId, &bd );
		switch ( i % 5 )
		{
			case 0:
			{
				b2Polygon box = b2MakeBox( pcg32_uniform( &rng, 0.05f, 0.3f ), pcg32_uniform( &rng, 0.05f, 0.3f ) );
				b2CreatePolygonShape( body, &sd, &box );
				break;
			}
			case 1:
			{
				b2Circle circle = { { 0.0f, 0.0f }, pcg32_uniform( &rng, 0.05f, 0.25f ) };
				b2CreateCircleShape( body, &sd, &circle );
				break;
			}
			case 2:
			{
				float h = pcg32_uniform( &rng, 0.1f, 0.4f );
				b2Capsule capsule = { { -h, 0.0f }, { h, 0.0f }, pcg32_uniform( &rng, 0.04f, 0.12f ) };
				b2CreateCapsuleShape( body, &sd, &capsule );
				break;
			}
			case 3:
			{
				b2Polygon rounded = b2MakeRoundedBox( pcg32_uniform( &rng, 0.05f, 0.2f ), pcg32_uniform( &rng, 0.05f, 0.2f ), 0.03f );
				b2CreatePolygonShape( body, &sd, &rounded );
				break;
			}
			default:
			{
				// two shapes: a box with a circle on one end
				b2Polygon box = b2MakeOffsetBox( 0.25f, 0.06f, ( b2Vec2 ){ 0.1f, 0.0f }, 0.0f );
				b2CreatePolygonShape( body, &sd, &box );
				b2Circle knob = { { -0.2f, 0.0f }, 0.1f };
				b2CreateCircleShape( body, &sd, &knob );
				break;
			}
		}
		if ( n < capacity )
			outBodies[n++] = body;
	}
	return n;
}
