// 2D vector math used by the public API and by every stage of the step pipeline.
//
// Types and semantics follow the reference's include/box2d/math_functions.h:18-566. The
// evaluation order of every inline below is part of the contract: the step pipeline is
// checked bit-for-bit against the reference, so e.g. b2MulAdd is a + s*b as two rounded
// float operations (never a fused multiply-add) and b2MinFloat is (a < b ? a : b).
#pragma once

#include "base.h"

#include <math.h>
#include <stdbool.h>

#define b2_pi 3.14159265359f

typedef struct b2Vec2
{
	float x, y;
} b2Vec2;

// cosine and sine of an angle
typedef struct b2Rot
{
	float c, s;
} b2Rot;

typedef struct b2Transform
{
	b2Vec2 p;
	b2Rot q;
} b2Transform;

typedef struct b2Mat22
{
	b2Vec2 cx, cy;
} b2Mat22;

typedef struct b2AABB
{
	b2Vec2 lowerBound;
	b2Vec2 upperBound;
} b2AABB;

static const b2Vec2 b2Vec2_zero = { 0.0f, 0.0f };
static const b2Rot b2Rot_identity = { 1.0f, 0.0f };
static const b2Transform b2Transform_identity = { { 0.0f, 0.0f }, { 1.0f, 0.0f } };

B2_INLINE float b2MinFloat( float a, float b ) { return a < b ? a : b; }
B2_INLINE float b2MaxFloat( float a, float b ) { return a > b ? a : b; }
B2_INLINE float b2AbsFloat( float a ) { return a < 0 ? -a : a; }
B2_INLINE float b2ClampFloat( float a, float lo, float hi ) { return a < lo ? lo : ( a > hi ? hi : a ); }
B2_INLINE int b2MinInt( int a, int b ) { return a < b ? a : b; }
B2_INLINE int b2MaxInt( int a, int b ) { return a > b ? a : b; }
B2_INLINE int b2AbsInt( int a ) { return a < 0 ? -a : a; }
B2_INLINE int b2ClampInt( int a, int lo, int hi ) { return a < lo ? lo : ( a > hi ? hi : a ); }

B2_INLINE float b2Dot( b2Vec2 a, b2Vec2 b ) { return a.x * b.x + a.y * b.y; }
B2_INLINE float b2Cross( b2Vec2 a, b2Vec2 b ) { return a.x * b.y - a.y * b.x; }
B2_INLINE b2Vec2 b2CrossVS( b2Vec2 v, float s ) { return B2_LITERAL( b2Vec2 ){ s * v.y, -s * v.x }; }
B2_INLINE b2Vec2 b2CrossSV( float s, b2Vec2 v ) { return B2_LITERAL( b2Vec2 ){ -s * v.y, s * v.x }; }
B2_INLINE b2Vec2 b2LeftPerp( b2Vec2 v ) { return B2_LITERAL( b2Vec2 ){ -v.y, v.x }; }
B2_INLINE b2Vec2 b2RightPerp( b2Vec2 v ) { return B2_LITERAL( b2Vec2 ){ v.y, -v.x }; }
B2_INLINE b2Vec2 b2Add( b2Vec2 a, b2Vec2 b ) { return B2_LITERAL( b2Vec2 ){ a.x + b.x, a.y + b.y }; }
B2_INLINE b2Vec2 b2Sub( b2Vec2 a, b2Vec2 b ) { return B2_LITERAL( b2Vec2 ){ a.x - b.x, a.y - b.y }; }
B2_INLINE b2Vec2 b2Neg( b2Vec2 a ) { return B2_LITERAL( b2Vec2 ){ -a.x, -a.y }; }
B2_INLINE b2Vec2 b2Lerp( b2Vec2 a, b2Vec2 b, float t )
{
	return B2_LITERAL( b2Vec2 ){ ( 1.0f - t ) * a.x + t * b.x, ( 1.0f - t ) * a.y + t * b.y };
}
B2_INLINE b2Vec2 b2Mul( b2Vec2 a, b2Vec2 b ) { return B2_LITERAL( b2Vec2 ){ a.x * b.x, a.y * b.y }; }
B2_INLINE b2Vec2 b2MulSV( float s, b2Vec2 v ) { return B2_LITERAL( b2Vec2 ){ s * v.x, s * v.y }; }
B2_INLINE b2Vec2 b2MulAdd( b2Vec2 a, float s, b2Vec2 b ) { return B2_LITERAL( b2Vec2 ){ a.x + s * b.x, a.y + s * b.y }; }
B2_INLINE b2Vec2 b2MulSub( b2Vec2 a, float s, b2Vec2 b ) { return B2_LITERAL( b2Vec2 ){ a.x - s * b.x, a.y - s * b.y }; }
B2_INLINE b2Vec2 b2Abs( b2Vec2 a ) { return B2_LITERAL( b2Vec2 ){ b2AbsFloat( a.x ), b2AbsFloat( a.y ) }; }
B2_INLINE b2Vec2 b2Min( b2Vec2 a, b2Vec2 b ) { return B2_LITERAL( b2Vec2 ){ b2MinFloat( a.x, b.x ), b2MinFloat( a.y, b.y ) }; }
B2_INLINE b2Vec2 b2Max( b2Vec2 a, b2Vec2 b ) { return B2_LITERAL( b2Vec2 ){ b2MaxFloat( a.x, b.x ), b2MaxFloat( a.y, b.y ) }; }
B2_INLINE float b2Length( b2Vec2 v ) { return sqrtf( v.x * v.x + v.y * v.y ); }
B2_INLINE float b2LengthSquared( b2Vec2 v ) { return v.x * v.x + v.y * v.y; }
B2_INLINE float b2Distance( b2Vec2 a, b2Vec2 b )
{
	float dx = b.x - a.x, dy = b.y - a.y;
	return sqrtf( dx * dx + dy * dy );
}
B2_INLINE float b2DistanceSquared( b2Vec2 a, b2Vec2 b )
{
	float dx = b.x - a.x, dy = b.y - a.y;
	return dx * dx + dy * dy;
}

B2_INLINE b2Rot b2MakeRot( float angle ) { return B2_LITERAL( b2Rot ){ cosf( angle ), sinf( angle ) }; }

B2_INLINE b2Rot b2NormalizeRot( b2Rot q )
{
	float mag = sqrtf( q.s * q.s + q.c * q.c );
	float inv = mag > 0.0f ? 1.0f / mag : 0.0f;
	return B2_LITERAL( b2Rot ){ q.c * inv, q.s * inv };
}

B2_INLINE bool b2IsNormalized( b2Rot q )
{
	float qq = q.s * q.s + q.c * q.c;
	return 1.0f - 0.0006f < qq && qq < 1.0f + 0.0006f;
}

B2_INLINE b2Rot b2NLerp( b2Rot q1, b2Rot q2, float t )
{
	float omt = 1.0f - t;
	b2Rot q = { omt * q1.c + t * q2.c, omt * q1.s + t * q2.s };
	return b2NormalizeRot( q );
}

// First-order rotation update followed by renormalization (reference
// include/box2d/math_functions.h:306-317).
B2_INLINE b2Rot b2IntegrateRotation( b2Rot q1, float deltaAngle )
{
	b2Rot q2 = { q1.c - deltaAngle * q1.s, q1.s + deltaAngle * q1.c };
	float mag = sqrtf( q2.s * q2.s + q2.c * q2.c );
	float inv = mag > 0.0f ? 1.0f / mag : 0.0f;
	return B2_LITERAL( b2Rot ){ q2.c * inv, q2.s * inv };
}

B2_INLINE float b2Rot_GetAngle( b2Rot q ) { return atan2f( q.s, q.c ); }
B2_INLINE b2Vec2 b2Rot_GetXAxis( b2Rot q ) { return B2_LITERAL( b2Vec2 ){ q.c, q.s }; }
B2_INLINE b2Vec2 b2Rot_GetYAxis( b2Rot q ) { return B2_LITERAL( b2Vec2 ){ -q.s, q.c }; }

B2_INLINE b2Rot b2MulRot( b2Rot q, b2Rot r )
{
	b2Rot qr;
	qr.s = q.s * r.c + q.c * r.s;
	qr.c = q.c * r.c - q.s * r.s;
	return qr;
}

B2_INLINE b2Rot b2InvMulRot( b2Rot q, b2Rot r )
{
	b2Rot qr;
	qr.s = q.c * r.s - q.s * r.c;
	qr.c = q.c * r.c + q.s * r.s;
	return qr;
}

B2_INLINE b2Vec2 b2RotateVector( b2Rot q, b2Vec2 v )
{
	return B2_LITERAL( b2Vec2 ){ q.c * v.x - q.s * v.y, q.s * v.x + q.c * v.y };
}

B2_INLINE b2Vec2 b2InvRotateVector( b2Rot q, b2Vec2 v )
{
	return B2_LITERAL( b2Vec2 ){ q.c * v.x + q.s * v.y, -q.s * v.x + q.c * v.y };
}

B2_INLINE b2Vec2 b2TransformPoint( b2Transform t, const b2Vec2 p )
{
	float x = ( t.q.c * p.x - t.q.s * p.y ) + t.p.x;
	float y = ( t.q.s * p.x + t.q.c * p.y ) + t.p.y;
	return B2_LITERAL( b2Vec2 ){ x, y };
}

B2_INLINE b2Vec2 b2InvTransformPoint( b2Transform t, const b2Vec2 p )
{
	float vx = p.x - t.p.x, vy = p.y - t.p.y;
	return B2_LITERAL( b2Vec2 ){ t.q.c * vx + t.q.s * vy, -t.q.s * vx + t.q.c * vy };
}

B2_INLINE b2Transform b2MulTransforms( b2Transform A, b2Transform B )
{
	b2Transform C;
	C.q = b2MulRot( A.q, B.q );
	C.p = b2Add( b2RotateVector( A.q, B.p ), A.p );
	return C;
}

B2_INLINE b2Transform b2InvMulTransforms( b2Transform A, b2Transform B )
{
	b2Transform C;
	C.q = b2InvMulRot( A.q, B.q );
	C.p = b2InvRotateVector( A.q, b2Sub( B.p, A.p ) );
	return C;
}

B2_INLINE bool b2AABB_Contains( b2AABB a, b2AABB b )
{
	return a.lowerBound.x <= b.lowerBound.x && a.lowerBound.y <= b.lowerBound.y && b.upperBound.x <= a.upperBound.x &&
		   b.upperBound.y <= a.upperBound.y;
}

B2_INLINE b2Vec2 b2AABB_Center( b2AABB a )
{
	return B2_LITERAL( b2Vec2 ){ 0.5f * ( a.lowerBound.x + a.upperBound.x ), 0.5f * ( a.lowerBound.y + a.upperBound.y ) };
}

B2_INLINE b2Vec2 b2AABB_Extents( b2AABB a )
{
	return B2_LITERAL( b2Vec2 ){ 0.5f * ( a.upperBound.x - a.lowerBound.x ), 0.5f * ( a.upperBound.y - a.lowerBound.y ) };
}

B2_INLINE b2AABB b2AABB_Union( b2AABB a, b2AABB b )
{
	b2AABB c;
	c.lowerBound.x = b2MinFloat( a.lowerBound.x, b.lowerBound.x );
	c.lowerBound.y = b2MinFloat( a.lowerBound.y, b.lowerBound.y );
	c.upperBound.x = b2MaxFloat( a.upperBound.x, b.upperBound.x );
	c.upperBound.y = b2MaxFloat( a.upperBound.y, b.upperBound.y );
	return c;
}

B2_API bool b2IsValid( float a );
B2_API bool b2Vec2_IsValid( b2Vec2 v );
B2_API bool b2Rot_IsValid( b2Rot q );
B2_API bool b2AABB_IsValid( b2AABB aabb );
B2_API b2Vec2 b2Normalize( b2Vec2 v );
B2_API b2Vec2 b2GetLengthAndNormalize( float* length, b2Vec2 v );
