// Joint solver of one step (prepare, warm start, solve/relax), shared by the step kernels'
// workgroup, cooperative-grid and per-phase paths. Each function restates the reference's joint
// function of the same stage in its operation order (src/joint.c:983-1164 dispatch; the per-type
// files cited below); with -ffp-contract=off and correctly rounded division and square root the
// results are the reference's floats.
//
// Body access goes through a solver store (solver.hpp): a negative local index is a static body,
// read as the identity state and never written, as the reference's dummy state
// (src/revolute_joint.c:283-284). Prepare reads the body sims: an item's body list for awake
// bodies, the static set for local index -1 - s (static sim s).
#pragma once

#include "solver.hpp"

#include "box2d/types.h"

namespace b2mi
{

// ---- arc tangent with the host C library's float results ---------------------------------
//
// b2RelativeAngle (include/box2d/math_functions.h:387-394) calls atan2f, so the joint stages
// depend on the bits of the host's atan2f. The reference's builds use glibc, whose float
// atan2f/atanf are Sun's fdlibm algorithm (e_atan2f.c, s_atanf.c: argument reduction to
// |x| < 7/16 about atan(0.5), atan(1), atan(1.5) or atan(inf) in hi/lo parts, then an 11-term
// odd polynomial split into even and odd halves). This is that published algorithm in float
// arithmetic, so it returns the same bits on the device; it is checked against the host's
// atan2f over tens of millions of arguments (tests/test_joints.py).

B2MI_HD inline uint32_t fbits( float f )
{
	return __builtin_bit_cast( uint32_t, f );
}

B2MI_HD inline float fdlibm_atanf( float x )
{
	// atan of 0.5, 1, 1.5 and +inf in a high part and a low correction
	const float atanhi[4] = { 4.6364760399e-01f, 7.8539812565e-01f, 9.8279368877e-01f, 1.5707962513e+00f };
	const float atanlo[4] = { 5.0121582440e-09f, 3.7748947079e-08f, 3.4473217170e-08f, 7.5497894159e-08f };
	const int32_t hx = (int32_t)fbits( x ), ix = hx & 0x7fffffff;
	int id;
	if ( ix >= 0x4c000000 ) // |x| >= 2^25
	{
		if ( ix > 0x7f800000 )
			return x + x; // NaN
		return hx > 0 ? atanhi[3] + atanlo[3] : -atanhi[3] - atanlo[3];
	}
	if ( ix < 0x3ee00000 ) // |x| < 7/16
	{
		if ( ix < 0x31000000 ) // |x| < 2^-29
			return x;
		id = -1;
	}
	else
	{
		x = fabsf( x );
		if ( ix < 0x3f980000 ) // |x| < 19/16
		{
			if ( ix < 0x3f300000 ) // 7/16 <= |x| < 11/16
			{
				id = 0;
				x = ( 2.0f * x - 1.0f ) / ( 2.0f + x );
			}
			else
			{
				id = 1;
				x = ( x - 1.0f ) / ( x + 1.0f );
			}
		}
		else if ( ix < 0x401c0000 ) // |x| < 39/16
		{
			id = 2;
			x = ( x - 1.5f ) / ( 1.0f + 1.5f * x );
		}
		else
		{
			id = 3;
			x = -1.0f / x;
		}
	}
	const float z = x * x;
	const float w = z * z;
	// odd and even halves of sum aT[i] z^(i+1), i = 0..10
	const float s1 = z * ( 3.3333334327e-01f +
						   w * ( 1.4285714924e-01f + w * ( 9.0908870101e-02f + w * ( 6.6610731184e-02f +
																						w * ( 4.9768779427e-02f + w * 1.6285819933e-02f ) ) ) ) );
	const float s2 = w * ( -2.0000000298e-01f +
						   w * ( -1.1111110449e-01f + w * ( -7.6918758452e-02f + w * ( -5.8335702866e-02f + w * -3.6531571299e-02f ) ) ) );
	if ( id < 0 )
		return x - x * ( s1 + s2 );
	const float r = atanhi[id] - ( ( x * ( s1 + s2 ) - atanlo[id] ) - x );
	return hx < 0 ? -r : r;
}

B2MI_HD inline float fdlibm_atan2f( float y, float x )
{
	const float pi_o_4 = 7.8539818525e-01f, pi_o_2 = 1.5707963705e+00f, pi = 3.1415927410e+00f;
	const float pi_lo = -8.7422776573e-08f, tiny = 1.0e-30f;
	const int32_t hx = (int32_t)fbits( x ), ix = hx & 0x7fffffff;
	const int32_t hy = (int32_t)fbits( y ), iy = hy & 0x7fffffff;
	if ( ix > 0x7f800000 || iy > 0x7f800000 )
		return x + y; // NaN
	if ( hx == 0x3f800000 )
		return fdlibm_atanf( y ); // x = 1
	const int m = ( ( hy >> 31 ) & 1 ) | ( ( hx >> 30 ) & 2 ); // 2 sign(x) + sign(y)
	if ( iy == 0 )
		return m <= 1 ? y : m == 2 ? pi + tiny : -pi - tiny;
	if ( ix == 0 )
		return hy < 0 ? -pi_o_2 - tiny : pi_o_2 + tiny;
	if ( ix == 0x7f800000 )
	{
		if ( iy == 0x7f800000 )
			return m == 0 ? pi_o_4 + tiny : m == 1 ? -pi_o_4 - tiny : m == 2 ? 3.0f * pi_o_4 + tiny : -3.0f * pi_o_4 - tiny;
		return m == 0 ? 0.0f : m == 1 ? -0.0f : m == 2 ? pi + tiny : -pi - tiny;
	}
	if ( iy == 0x7f800000 )
		return hy < 0 ? -pi_o_2 - tiny : pi_o_2 + tiny;
	const int k = ( iy - ix ) >> 23;
	float z;
	if ( k > 26 )
		z = pi_o_2 + 0.5f * pi_lo; // |y/x| > 2^26
	else if ( hx < 0 && k < -26 )
		z = 0.0f; // |y|/x < -2^26
	else
		z = fdlibm_atanf( fabsf( y / x ) );
	switch ( m )
	{
		case 0:
			return z;
		case 1:
			return -z;
		case 2:
			return pi - ( z - pi_lo );
		default:
			return ( z - pi_lo ) - pi;
	}
}

// include/box2d/math_functions.h:387-394 (b2RelativeAngle), 397-409 (b2UnwindAngle),
// 489-499 (b2Solve22)
B2MI_HD inline float relative_angle( b2Rot b, b2Rot a )
{
	float s = b.s * a.c - b.c * a.s;
	float c = b.c * a.c + b.s * a.s;
	return fdlibm_atan2f( s, c );
}

B2MI_HD inline float unwind_angle( float angle )
{
	if ( angle < -b2_pi )
		return angle + 2.0f * b2_pi;
	if ( angle > b2_pi )
		return angle - 2.0f * b2_pi;
	return angle;
}

B2MI_HD inline b2Vec2 solve22( b2Mat22 A, b2Vec2 b )
{
	float a11 = A.cx.x, a12 = A.cy.x, a21 = A.cx.y, a22 = A.cy.y;
	float det = a11 * a22 - a12 * a21;
	if ( det != 0.0f )
		det = 1.0f / det;
	return b2Vec2{ det * ( a22 * b.x - a12 * b.y ), det * ( a11 * b.y - a21 * b.x ) };
}

// src/solver.h:144-156
B2MI_HD inline DevSoft make_soft( float hertz, float zeta, float h )
{
	if ( hertz == 0.0f )
		return DevSoft{ 0.0f, 1.0f, 0.0f };
	float omega = 2.0f * b2_pi * hertz;
	float a1 = 2.0f * zeta + h * omega;
	float a2 = h * omega * a1;
	float a3 = 1.0f / ( 1.0f + a2 );
	return DevSoft{ omega / a1, a2 * a3, a3 };
}

__device__ inline DevSoft joint_softness( const DevWorld& W )
{
	return DevSoft{ W.jointSoft[0], W.jointSoft[1], W.jointSoft[2] };
}

// the sim a joint's prepare reads for local body li of its item
__device__ inline const DevBodySim& joint_body_sim( const DevWorld& W, const int* __restrict__ bl, int li )
{
	return li >= 0 ? W.sims[bl[li]] : W.statics[-1 - li];
}

__device__ inline b2Vec2 sim_center( const DevBodySim& s )
{
	return b2Vec2{ s.center[0], s.center[1] };
}

__device__ inline b2Vec2 sim_local_center( const DevBodySim& s )
{
	return b2Vec2{ s.localCenter[0], s.localCenter[1] };
}

__device__ inline b2Rot sim_rot( const DevBodySim& s )
{
	return b2Rot{ s.q[0], s.q[1] };
}

__device__ inline b2Rot body_dq( const Body& b )
{
	return b2Rot{ b.dqc, b.dqs };
}

// ---- revolute (src/revolute_joint.c:201-467) ------------------------------------------------

__device__ inline void prepare_revolute( DevJoint& base, const DevBodySim& simA, const DevBodySim& simB, const DevWorld& W )
{
	DevRevoluteJoint& joint = base.revolute;
	float iA = simA.invI, iB = simB.invI;
	joint.anchorA = b2RotateVector( sim_rot( simA ), b2Sub( base.localOriginAnchorA, sim_local_center( simA ) ) );
	joint.anchorB = b2RotateVector( sim_rot( simB ), b2Sub( base.localOriginAnchorB, sim_local_center( simB ) ) );
	joint.deltaCenter = b2Sub( sim_center( simB ), sim_center( simA ) );
	joint.deltaAngle = relative_angle( sim_rot( simB ), sim_rot( simA ) ) - joint.referenceAngle;
	joint.deltaAngle = unwind_angle( joint.deltaAngle );
	float k = iA + iB;
	joint.axialMass = k > 0.0f ? 1.0f / k : 0.0f;
	joint.springSoftness = make_soft( joint.hertz, joint.dampingRatio, W.h );
	if ( ( W.flags & kWorldWarmStart ) == 0 )
	{
		joint.linearImpulse = b2Vec2{ 0.0f, 0.0f };
		joint.springImpulse = 0.0f;
		joint.motorImpulse = 0.0f;
		joint.lowerImpulse = 0.0f;
		joint.upperImpulse = 0.0f;
	}
}

__device__ inline void warm_start_revolute( const DevJoint& base, Body& A, Body& B )
{
	const DevRevoluteJoint& joint = base.revolute;
	float mA = base.invMassA, mB = base.invMassB, iA = base.invIA, iB = base.invIB;
	b2Vec2 rA = b2RotateVector( body_dq( A ), joint.anchorA );
	b2Vec2 rB = b2RotateVector( body_dq( B ), joint.anchorB );
	float axialImpulse = joint.springImpulse + joint.motorImpulse + joint.lowerImpulse - joint.upperImpulse;
	A.vx = A.vx - mA * joint.linearImpulse.x;
	A.vy = A.vy - mA * joint.linearImpulse.y;
	A.w -= iA * ( b2Cross( rA, joint.linearImpulse ) + axialImpulse );
	B.vx = B.vx + mB * joint.linearImpulse.x;
	B.vy = B.vy + mB * joint.linearImpulse.y;
	B.w += iB * ( b2Cross( rB, joint.linearImpulse ) + axialImpulse );
}

__device__ inline void solve_revolute( DevJoint& base, Body& A, Body& B, const DevWorld& W, bool useBias )
{
	DevRevoluteJoint& joint = base.revolute;
	float mA = base.invMassA, mB = base.invMassB, iA = base.invIA, iB = base.invIB;
	b2Vec2 vA = { A.vx, A.vy }, vB = { B.vx, B.vy };
	float wA = A.w, wB = B.w;
	const bool fixedRotation = ( iA + iB == 0.0f );
	const DevSoft jointSoftness = joint_softness( W );

	if ( ( base.flags & kJointSpring ) && fixedRotation == false )
	{
		float C = relative_angle( body_dq( B ), body_dq( A ) ) + joint.deltaAngle;
		float bias = joint.springSoftness.biasRate * C;
		float massScale = joint.springSoftness.massScale;
		float impulseScale = joint.springSoftness.impulseScale;
		float Cdot = wB - wA;
		float impulse = -massScale * joint.axialMass * ( Cdot + bias ) - impulseScale * joint.springImpulse;
		joint.springImpulse += impulse;
		wA -= iA * impulse;
		wB += iB * impulse;
	}

	if ( ( base.flags & kJointMotor ) && fixedRotation == false )
	{
		float Cdot = wB - wA - joint.motorSpeed;
		float impulse = -joint.axialMass * Cdot;
		float oldImpulse = joint.motorImpulse;
		float maxImpulse = W.h * joint.maxMotorTorque;
		joint.motorImpulse = b2ClampFloat( joint.motorImpulse + impulse, -maxImpulse, maxImpulse );
		impulse = joint.motorImpulse - oldImpulse;
		wA -= iA * impulse;
		wB += iB * impulse;
	}

	if ( ( base.flags & kJointLimit ) && fixedRotation == false )
	{
		float jointAngle = relative_angle( body_dq( B ), body_dq( A ) ) + joint.deltaAngle;
		jointAngle = unwind_angle( jointAngle );
		// lower limit
		{
			float C = jointAngle - joint.lowerAngle;
			float bias = 0.0f, massScale = 1.0f, impulseScale = 0.0f;
			if ( C > 0.0f )
				bias = C * W.inv_h; // speculation
			else if ( useBias )
			{
				bias = jointSoftness.biasRate * C;
				massScale = jointSoftness.massScale;
				impulseScale = jointSoftness.impulseScale;
			}
			float Cdot = wB - wA;
			float impulse = -massScale * joint.axialMass * ( Cdot + bias ) - impulseScale * joint.lowerImpulse;
			float oldImpulse = joint.lowerImpulse;
			joint.lowerImpulse = b2MaxFloat( joint.lowerImpulse + impulse, 0.0f );
			impulse = joint.lowerImpulse - oldImpulse;
			wA -= iA * impulse;
			wB += iB * impulse;
		}
		// upper limit (signs flipped; the reference scales the lower impulse here,
		// src/revolute_joint.c:405)
		{
			float C = joint.upperAngle - jointAngle;
			float bias = 0.0f, massScale = 1.0f, impulseScale = 0.0f;
			if ( C > 0.0f )
				bias = C * W.inv_h;
			else if ( useBias )
			{
				bias = jointSoftness.biasRate * C;
				massScale = jointSoftness.massScale;
				impulseScale = jointSoftness.impulseScale;
			}
			float Cdot = wA - wB;
			float impulse = -massScale * joint.axialMass * ( Cdot + bias ) - impulseScale * joint.lowerImpulse;
			float oldImpulse = joint.upperImpulse;
			joint.upperImpulse = b2MaxFloat( joint.upperImpulse + impulse, 0.0f );
			impulse = joint.upperImpulse - oldImpulse;
			wA += iA * impulse;
			wB -= iB * impulse;
		}
	}

	// point-to-point constraint
	{
		b2Vec2 rA = b2RotateVector( body_dq( A ), joint.anchorA );
		b2Vec2 rB = b2RotateVector( body_dq( B ), joint.anchorB );
		b2Vec2 Cdot = b2Sub( b2Add( vB, b2CrossSV( wB, rB ) ), b2Add( vA, b2CrossSV( wA, rA ) ) );
		b2Vec2 bias = { 0.0f, 0.0f };
		float massScale = 1.0f, impulseScale = 0.0f;
		if ( useBias )
		{
			b2Vec2 dcA = { A.dpx, A.dpy }, dcB = { B.dpx, B.dpy };
			b2Vec2 separation = b2Add( b2Add( b2Sub( dcB, dcA ), b2Sub( rB, rA ) ), joint.deltaCenter );
			bias = b2MulSV( jointSoftness.biasRate, separation );
			massScale = jointSoftness.massScale;
			impulseScale = jointSoftness.impulseScale;
		}
		b2Mat22 K;
		K.cx.x = mA + mB + rA.y * rA.y * iA + rB.y * rB.y * iB;
		K.cy.x = -rA.y * rA.x * iA - rB.y * rB.x * iB;
		K.cx.y = K.cy.x;
		K.cy.y = mA + mB + rA.x * rA.x * iA + rB.x * rB.x * iB;
		b2Vec2 b = solve22( K, b2Add( Cdot, bias ) );
		b2Vec2 impulse;
		impulse.x = -massScale * b.x - impulseScale * joint.linearImpulse.x;
		impulse.y = -massScale * b.y - impulseScale * joint.linearImpulse.y;
		joint.linearImpulse.x += impulse.x;
		joint.linearImpulse.y += impulse.y;
		vA = b2MulSub( vA, mA, impulse );
		wA -= iA * b2Cross( rA, impulse );
		vB = b2MulAdd( vB, mB, impulse );
		wB += iB * b2Cross( rB, impulse );
	}
	A.vx = vA.x;
	A.vy = vA.y;
	A.w = wA;
	B.vx = vB.x;
	B.vy = vB.y;
	B.w = wB;
}

// ---- weld (src/weld_joint.c:85-297) ---------------------------------------------------------

__device__ inline void prepare_weld( DevJoint& base, const DevBodySim& simA, const DevBodySim& simB, const DevWorld& W )
{
	DevWeldJoint& joint = base.weld;
	float iA = simA.invI, iB = simB.invI;
	b2Rot qA = sim_rot( simA ), qB = sim_rot( simB );
	joint.anchorA = b2RotateVector( qA, b2Sub( base.localOriginAnchorA, sim_local_center( simA ) ) );
	joint.anchorB = b2RotateVector( qB, b2Sub( base.localOriginAnchorB, sim_local_center( simB ) ) );
	joint.deltaCenter = b2Sub( sim_center( simB ), sim_center( simA ) );
	joint.deltaAngle = relative_angle( qB, qA ) - joint.referenceAngle;
	float ka = iA + iB;
	joint.axialMass = ka > 0.0f ? 1.0f / ka : 0.0f;
	joint.linearSoftness =
		joint.linearHertz == 0.0f ? joint_softness( W ) : make_soft( joint.linearHertz, joint.linearDampingRatio, W.h );
	joint.angularSoftness =
		joint.angularHertz == 0.0f ? joint_softness( W ) : make_soft( joint.angularHertz, joint.angularDampingRatio, W.h );
	if ( ( W.flags & kWorldWarmStart ) == 0 )
	{
		joint.linearImpulse = b2Vec2{ 0.0f, 0.0f };
		joint.angularImpulse = 0.0f;
	}
}

__device__ inline void warm_start_weld( const DevJoint& base, Body& A, Body& B )
{
	const DevWeldJoint& joint = base.weld;
	float mA = base.invMassA, mB = base.invMassB, iA = base.invIA, iB = base.invIB;
	b2Vec2 rA = b2RotateVector( body_dq( A ), joint.anchorA );
	b2Vec2 rB = b2RotateVector( body_dq( B ), joint.anchorB );
	A.vx = A.vx - mA * joint.linearImpulse.x;
	A.vy = A.vy - mA * joint.linearImpulse.y;
	A.w -= iA * ( b2Cross( rA, joint.linearImpulse ) + joint.angularImpulse );
	B.vx = B.vx + mB * joint.linearImpulse.x;
	B.vy = B.vy + mB * joint.linearImpulse.y;
	B.w += iB * ( b2Cross( rB, joint.linearImpulse ) + joint.angularImpulse );
}

__device__ inline void solve_weld( DevJoint& base, Body& A, Body& B, bool useBias )
{
	DevWeldJoint& joint = base.weld;
	float mA = base.invMassA, mB = base.invMassB, iA = base.invIA, iB = base.invIB;
	b2Vec2 vA = { A.vx, A.vy }, vB = { B.vx, B.vy };
	float wA = A.w, wB = B.w;
	// angular constraint
	{
		float bias = 0.0f, massScale = 1.0f, impulseScale = 0.0f;
		if ( useBias || joint.angularHertz > 0.0f )
		{
			float C = relative_angle( body_dq( B ), body_dq( A ) ) + joint.deltaAngle;
			bias = joint.angularSoftness.biasRate * C;
			massScale = joint.angularSoftness.massScale;
			impulseScale = joint.angularSoftness.impulseScale;
		}
		float Cdot = wB - wA;
		float impulse = -massScale * joint.axialMass * ( Cdot + bias ) - impulseScale * joint.angularImpulse;
		joint.angularImpulse += impulse;
		wA -= iA * impulse;
		wB += iB * impulse;
	}
	// linear constraint
	{
		b2Vec2 rA = b2RotateVector( body_dq( A ), joint.anchorA );
		b2Vec2 rB = b2RotateVector( body_dq( B ), joint.anchorB );
		b2Vec2 bias = { 0.0f, 0.0f };
		float massScale = 1.0f, impulseScale = 0.0f;
		if ( useBias || joint.linearHertz > 0.0f )
		{
			b2Vec2 dcA = { A.dpx, A.dpy }, dcB = { B.dpx, B.dpy };
			b2Vec2 C = b2Add( b2Add( b2Sub( dcB, dcA ), b2Sub( rB, rA ) ), joint.deltaCenter );
			bias = b2MulSV( joint.linearSoftness.biasRate, C );
			massScale = joint.linearSoftness.massScale;
			impulseScale = joint.linearSoftness.impulseScale;
		}
		b2Vec2 Cdot = b2Sub( b2Add( vB, b2CrossSV( wB, rB ) ), b2Add( vA, b2CrossSV( wA, rA ) ) );
		b2Mat22 K;
		K.cx.x = mA + mB + rA.y * rA.y * iA + rB.y * rB.y * iB;
		K.cy.x = -rA.y * rA.x * iA - rB.y * rB.x * iB;
		K.cx.y = K.cy.x;
		K.cy.y = mA + mB + rA.x * rA.x * iA + rB.x * rB.x * iB;
		b2Vec2 b = solve22( K, b2Add( Cdot, bias ) );
		b2Vec2 impulse = { -massScale * b.x - impulseScale * joint.linearImpulse.x,
						   -massScale * b.y - impulseScale * joint.linearImpulse.y };
		joint.linearImpulse = b2Add( joint.linearImpulse, impulse );
		vA = b2MulSub( vA, mA, impulse );
		wA -= iA * b2Cross( rA, impulse );
		vB = b2MulAdd( vB, mB, impulse );
		wB += iB * b2Cross( rB, impulse );
	}
	A.vx = vA.x;
	A.vy = vA.y;
	A.w = wA;
	B.vx = vB.x;
	B.vy = vB.y;
	B.w = wB;
}

// ---- helpers shared by the axis joints --------------------------------------------------------

// b2Normalize (src/math_functions.c:55-66) and b2Dot, b2LeftPerp, b2MulMV, b2GetInverse22
// (include/box2d/math_functions.h:57, 146-150, 461-485)
__device__ inline b2Vec2 normalize( b2Vec2 v )
{
	return d_normalize( v );
}

__device__ inline b2Mat22 inverse22( b2Mat22 A )
{
	float a = A.cx.x, b = A.cy.x, c = A.cx.y, d = A.cy.y;
	float det = a * d - b * c;
	if ( det != 0.0f )
		det = 1.0f / det;
	b2Mat22 B;
	B.cx = b2Vec2{ det * d, -det * c };
	B.cy = b2Vec2{ -det * b, det * a };
	return B;
}

__device__ inline b2Vec2 mul_mv( b2Mat22 A, b2Vec2 v )
{
	return b2Vec2{ A.cx.x * v.x + A.cy.x * v.y, A.cx.y * v.x + A.cy.y * v.y };
}

// applies a linear impulse P with angular parts LA, LB: vA -= mA P, wA -= iA LA, vB += mB P, wB += iB LB
__device__ inline void apply_axis( b2Vec2& vA, float& wA, b2Vec2& vB, float& wB, float mA, float iA, float mB, float iB, b2Vec2 P, float LA,
								   float LB )
{
	vA = b2MulSub( vA, mA, P );
	wA -= iA * LA;
	vB = b2MulAdd( vB, mB, P );
	wB += iB * LB;
}

__device__ inline void store_bodies( Body& A, Body& B, b2Vec2 vA, float wA, b2Vec2 vB, float wB )
{
	A.vx = vA.x;
	A.vy = vA.y;
	A.w = wA;
	B.vx = vB.x;
	B.vy = vB.y;
	B.w = wB;
}

// ---- distance (src/distance_joint.c:196-461) ------------------------------------------------

__device__ inline void prepare_distance( DevJoint& base, const DevBodySim& simA, const DevBodySim& simB, const DevWorld& W )
{
	DevDistanceJoint& joint = base.distance;
	float mA = simA.invMass, iA = simA.invI, mB = simB.invMass, iB = simB.invI;
	joint.anchorA = b2RotateVector( sim_rot( simA ), b2Sub( base.localOriginAnchorA, sim_local_center( simA ) ) );
	joint.anchorB = b2RotateVector( sim_rot( simB ), b2Sub( base.localOriginAnchorB, sim_local_center( simB ) ) );
	joint.deltaCenter = b2Sub( sim_center( simB ), sim_center( simA ) );
	b2Vec2 rA = joint.anchorA, rB = joint.anchorB;
	b2Vec2 separation = b2Add( b2Sub( rB, rA ), joint.deltaCenter );
	b2Vec2 axis = normalize( separation );
	float crA = b2Cross( rA, axis );
	float crB = b2Cross( rB, axis );
	float k = mA + mB + iA * crA * crA + iB * crB * crB;
	joint.axialMass = k > 0.0f ? 1.0f / k : 0.0f;
	joint.distanceSoftness = make_soft( joint.hertz, joint.dampingRatio, W.h );
	if ( ( W.flags & kWorldWarmStart ) == 0 )
	{
		joint.impulse = 0.0f;
		joint.lowerImpulse = 0.0f;
		joint.upperImpulse = 0.0f;
		joint.motorImpulse = 0.0f;
	}
}

__device__ inline void warm_start_distance( const DevJoint& base, Body& A, Body& B )
{
	const DevDistanceJoint& joint = base.distance;
	float mA = base.invMassA, mB = base.invMassB, iA = base.invIA, iB = base.invIB;
	b2Vec2 rA = b2RotateVector( body_dq( A ), joint.anchorA );
	b2Vec2 rB = b2RotateVector( body_dq( B ), joint.anchorB );
	b2Vec2 ds = b2Add( b2Sub( b2Vec2{ B.dpx, B.dpy }, b2Vec2{ A.dpx, A.dpy } ), b2Sub( rB, rA ) );
	b2Vec2 separation = b2Add( joint.deltaCenter, ds );
	b2Vec2 axis = normalize( separation );
	float axialImpulse = joint.impulse + joint.lowerImpulse - joint.upperImpulse + joint.motorImpulse;
	b2Vec2 P = b2MulSV( axialImpulse, axis );
	A.vx = A.vx - mA * P.x;
	A.vy = A.vy - mA * P.y;
	A.w -= iA * b2Cross( rA, P );
	B.vx = B.vx + mB * P.x;
	B.vy = B.vy + mB * P.y;
	B.w += iB * b2Cross( rB, P );
}

// relative velocity of B's anchor against A's: (vB - vA) + (wB x rB - wA x rA)
__device__ inline b2Vec2 anchor_velocity( b2Vec2 vA, float wA, b2Vec2 rA, b2Vec2 vB, float wB, b2Vec2 rB )
{
	return b2Add( b2Sub( vB, vA ), b2Sub( b2CrossSV( wB, rB ), b2CrossSV( wA, rA ) ) );
}

__device__ inline void solve_distance( DevJoint& base, Body& A, Body& B, const DevWorld& W, bool useBias )
{
	DevDistanceJoint& joint = base.distance;
	float mA = base.invMassA, mB = base.invMassB, iA = base.invIA, iB = base.invIB;
	b2Vec2 vA = { A.vx, A.vy }, vB = { B.vx, B.vy };
	float wA = A.w, wB = B.w;
	const DevSoft jointSoftness = joint_softness( W );
	b2Vec2 rA = b2RotateVector( body_dq( A ), joint.anchorA );
	b2Vec2 rB = b2RotateVector( body_dq( B ), joint.anchorB );
	b2Vec2 ds = b2Add( b2Sub( b2Vec2{ B.dpx, B.dpy }, b2Vec2{ A.dpx, A.dpy } ), b2Sub( rB, rA ) );
	b2Vec2 separation = b2Add( joint.deltaCenter, ds );
	float length = b2Length( separation );
	b2Vec2 axis = normalize( separation );
	const bool enableSpring = ( base.flags & kJointSpring ) != 0, enableLimit = ( base.flags & kJointLimit ) != 0;
	if ( enableSpring && ( joint.minLength < joint.maxLength || enableLimit == false ) )
	{
		if ( joint.hertz > 0.0f )
		{
			b2Vec2 vr = anchor_velocity( vA, wA, rA, vB, wB, rB );
			float Cdot = b2Dot( axis, vr );
			float C = length - joint.length;
			float bias = joint.distanceSoftness.biasRate * C;
			float m = joint.distanceSoftness.massScale * joint.axialMass;
			float impulse = -m * ( Cdot + bias ) - joint.distanceSoftness.impulseScale * joint.impulse;
			joint.impulse += impulse;
			b2Vec2 P = b2MulSV( impulse, axis );
			apply_axis( vA, wA, vB, wB, mA, iA, mB, iB, P, b2Cross( rA, P ), b2Cross( rB, P ) );
		}
		if ( enableLimit )
		{
			// lower
			{
				b2Vec2 vr = anchor_velocity( vA, wA, rA, vB, wB, rB );
				float Cdot = b2Dot( axis, vr );
				float C = length - joint.minLength;
				float bias = 0.0f, massCoeff = 1.0f, impulseCoeff = 0.0f;
				if ( C > 0.0f )
					bias = C * W.inv_h;
				else if ( useBias )
				{
					bias = jointSoftness.biasRate * C;
					massCoeff = jointSoftness.massScale;
					impulseCoeff = jointSoftness.impulseScale;
				}
				float impulse = -massCoeff * joint.axialMass * ( Cdot + bias ) - impulseCoeff * joint.lowerImpulse;
				float newImpulse = b2MaxFloat( 0.0f, joint.lowerImpulse + impulse );
				impulse = newImpulse - joint.lowerImpulse;
				joint.lowerImpulse = newImpulse;
				b2Vec2 P = b2MulSV( impulse, axis );
				apply_axis( vA, wA, vB, wB, mA, iA, mB, iB, P, b2Cross( rA, P ), b2Cross( rB, P ) );
			}
			// upper
			{
				b2Vec2 vr = b2Add( b2Sub( vA, vB ), b2Sub( b2CrossSV( wA, rA ), b2CrossSV( wB, rB ) ) );
				float Cdot = b2Dot( axis, vr );
				float C = joint.maxLength - length;
				float bias = 0.0f, massScale = 1.0f, impulseScale = 0.0f;
				if ( C > 0.0f )
					bias = C * W.inv_h;
				else if ( useBias )
				{
					bias = jointSoftness.biasRate * C;
					massScale = jointSoftness.massScale;
					impulseScale = jointSoftness.impulseScale;
				}
				float impulse = -massScale * joint.axialMass * ( Cdot + bias ) - impulseScale * joint.upperImpulse;
				float newImpulse = b2MaxFloat( 0.0f, joint.upperImpulse + impulse );
				impulse = newImpulse - joint.upperImpulse;
				joint.upperImpulse = newImpulse;
				b2Vec2 P = b2MulSV( -impulse, axis );
				apply_axis( vA, wA, vB, wB, mA, iA, mB, iB, P, b2Cross( rA, P ), b2Cross( rB, P ) );
			}
		}
		if ( base.flags & kJointMotor )
		{
			b2Vec2 vr = anchor_velocity( vA, wA, rA, vB, wB, rB );
			float Cdot = b2Dot( axis, vr );
			float impulse = joint.axialMass * ( joint.motorSpeed - Cdot );
			float oldImpulse = joint.motorImpulse;
			float maxImpulse = W.h * joint.maxMotorForce;
			joint.motorImpulse = b2ClampFloat( joint.motorImpulse + impulse, -maxImpulse, maxImpulse );
			impulse = joint.motorImpulse - oldImpulse;
			b2Vec2 P = b2MulSV( impulse, axis );
			apply_axis( vA, wA, vB, wB, mA, iA, mB, iB, P, b2Cross( rA, P ), b2Cross( rB, P ) );
		}
	}
	else
	{
		// rigid constraint
		b2Vec2 vr = anchor_velocity( vA, wA, rA, vB, wB, rB );
		float Cdot = b2Dot( axis, vr );
		float C = length - joint.length;
		float bias = 0.0f, massScale = 1.0f, impulseScale = 0.0f;
		if ( useBias )
		{
			bias = jointSoftness.biasRate * C;
			massScale = jointSoftness.massScale;
			impulseScale = jointSoftness.impulseScale;
		}
		float impulse = -massScale * joint.axialMass * ( Cdot + bias ) - impulseScale * joint.impulse;
		joint.impulse += impulse;
		b2Vec2 P = b2MulSV( impulse, axis );
		apply_axis( vA, wA, vB, wB, mA, iA, mB, iB, P, b2Cross( rA, P ), b2Cross( rB, P ) );
	}
	store_bodies( A, B, vA, wA, vB, wB );
}

// ---- prismatic (src/prismatic_joint.c:203-480) ----------------------------------------------

__device__ inline void prepare_prismatic( DevJoint& base, const DevBodySim& simA, const DevBodySim& simB, const DevWorld& W )
{
	DevPrismaticJoint& joint = base.prismatic;
	float mA = simA.invMass, iA = simA.invI, mB = simB.invMass, iB = simB.invI;
	b2Rot qA = sim_rot( simA ), qB = sim_rot( simB );
	joint.anchorA = b2RotateVector( qA, b2Sub( base.localOriginAnchorA, sim_local_center( simA ) ) );
	joint.anchorB = b2RotateVector( qB, b2Sub( base.localOriginAnchorB, sim_local_center( simB ) ) );
	joint.axisA = b2RotateVector( qA, joint.localAxisA );
	joint.deltaCenter = b2Sub( sim_center( simB ), sim_center( simA ) );
	joint.deltaAngle = relative_angle( qB, qA ) - joint.referenceAngle;
	b2Vec2 rA = joint.anchorA, rB = joint.anchorB;
	b2Vec2 d = b2Add( joint.deltaCenter, b2Sub( rB, rA ) );
	float a1 = b2Cross( b2Add( d, rA ), joint.axisA );
	float a2 = b2Cross( rB, joint.axisA );
	float k = mA + mB + iA * a1 * a1 + iB * a2 * a2;
	joint.axialMass = k > 0.0f ? 1.0f / k : 0.0f;
	joint.springSoftness = make_soft( joint.hertz, joint.dampingRatio, W.h );
	if ( ( W.flags & kWorldWarmStart ) == 0 )
	{
		joint.impulse = b2Vec2{ 0.0f, 0.0f };
		joint.springImpulse = 0.0f;
		joint.motorImpulse = 0.0f;
		joint.lowerImpulse = 0.0f;
		joint.upperImpulse = 0.0f;
	}
}

// the separation of the anchors along the current frame: ((dpB - dpA) + deltaCenter) + (rB - rA)
__device__ inline b2Vec2 axis_separation( const Body& A, const Body& B, b2Vec2 deltaCenter, b2Vec2 rA, b2Vec2 rB )
{
	return b2Add( b2Add( b2Sub( b2Vec2{ B.dpx, B.dpy }, b2Vec2{ A.dpx, A.dpy } ), deltaCenter ), b2Sub( rB, rA ) );
}

__device__ inline void warm_start_prismatic( const DevJoint& base, Body& A, Body& B )
{
	const DevPrismaticJoint& joint = base.prismatic;
	float mA = base.invMassA, mB = base.invMassB, iA = base.invIA, iB = base.invIB;
	b2Vec2 rA = b2RotateVector( body_dq( A ), joint.anchorA );
	b2Vec2 rB = b2RotateVector( body_dq( B ), joint.anchorB );
	b2Vec2 d = axis_separation( A, B, joint.deltaCenter, rA, rB );
	b2Vec2 axisA = b2RotateVector( body_dq( A ), joint.axisA );
	float a1 = b2Cross( b2Add( d, rA ), axisA );
	float a2 = b2Cross( rB, axisA );
	float axialImpulse = joint.springImpulse + joint.motorImpulse + joint.lowerImpulse - joint.upperImpulse;
	b2Vec2 perpA = b2LeftPerp( axisA );
	float s1 = b2Cross( b2Add( d, rA ), perpA );
	float s2 = b2Cross( rB, perpA );
	float perpImpulse = joint.impulse.x;
	float angleImpulse = joint.impulse.y;
	b2Vec2 P = b2Add( b2MulSV( axialImpulse, axisA ), b2MulSV( perpImpulse, perpA ) );
	float LA = axialImpulse * a1 + perpImpulse * s1 + angleImpulse;
	float LB = axialImpulse * a2 + perpImpulse * s2 + angleImpulse;
	A.vx = A.vx - mA * P.x;
	A.vy = A.vy - mA * P.y;
	A.w -= iA * LA;
	B.vx = B.vx + mB * P.x;
	B.vy = B.vy + mB * P.y;
	B.w += iB * LB;
}

__device__ inline void solve_prismatic( DevJoint& base, Body& A, Body& B, const DevWorld& W, bool useBias )
{
	DevPrismaticJoint& joint = base.prismatic;
	float mA = base.invMassA, mB = base.invMassB, iA = base.invIA, iB = base.invIB;
	b2Vec2 vA = { A.vx, A.vy }, vB = { B.vx, B.vy };
	float wA = A.w, wB = B.w;
	const DevSoft jointSoftness = joint_softness( W );
	b2Vec2 rA = b2RotateVector( body_dq( A ), joint.anchorA );
	b2Vec2 rB = b2RotateVector( body_dq( B ), joint.anchorB );
	b2Vec2 d = axis_separation( A, B, joint.deltaCenter, rA, rB );
	b2Vec2 axisA = b2RotateVector( body_dq( A ), joint.axisA );
	float translation = b2Dot( axisA, d );
	float a1 = b2Cross( b2Add( d, rA ), axisA );
	float a2 = b2Cross( rB, axisA );
	if ( base.flags & kJointSpring )
	{
		float C = translation;
		float bias = joint.springSoftness.biasRate * C;
		float massScale = joint.springSoftness.massScale;
		float impulseScale = joint.springSoftness.impulseScale;
		float Cdot = b2Dot( axisA, b2Sub( vB, vA ) ) + a2 * wB - a1 * wA;
		float impulse = -massScale * joint.axialMass * ( Cdot + bias ) - impulseScale * joint.springImpulse;
		joint.springImpulse += impulse;
		apply_axis( vA, wA, vB, wB, mA, iA, mB, iB, b2MulSV( impulse, axisA ), impulse * a1, impulse * a2 );
	}
	if ( base.flags & kJointMotor )
	{
		float Cdot = b2Dot( axisA, b2Sub( vB, vA ) ) + a2 * wB - a1 * wA;
		float impulse = joint.axialMass * ( joint.motorSpeed - Cdot );
		float oldImpulse = joint.motorImpulse;
		float maxImpulse = W.h * joint.maxMotorForce;
		joint.motorImpulse = b2ClampFloat( joint.motorImpulse + impulse, -maxImpulse, maxImpulse );
		impulse = joint.motorImpulse - oldImpulse;
		apply_axis( vA, wA, vB, wB, mA, iA, mB, iB, b2MulSV( impulse, axisA ), impulse * a1, impulse * a2 );
	}
	if ( base.flags & kJointLimit )
	{
		// lower
		{
			float C = translation - joint.lowerTranslation;
			float bias = 0.0f, massScale = 1.0f, impulseScale = 0.0f;
			if ( C > 0.0f )
				bias = C * W.inv_h;
			else if ( useBias )
			{
				bias = jointSoftness.biasRate * C;
				massScale = jointSoftness.massScale;
				impulseScale = jointSoftness.impulseScale;
			}
			float oldImpulse = joint.lowerImpulse;
			float Cdot = b2Dot( axisA, b2Sub( vB, vA ) ) + a2 * wB - a1 * wA;
			float impulse = -joint.axialMass * massScale * ( Cdot + bias ) - impulseScale * oldImpulse;
			joint.lowerImpulse = b2MaxFloat( oldImpulse + impulse, 0.0f );
			impulse = joint.lowerImpulse - oldImpulse;
			apply_axis( vA, wA, vB, wB, mA, iA, mB, iB, b2MulSV( impulse, axisA ), impulse * a1, impulse * a2 );
		}
		// upper (signs flipped)
		{
			float C = joint.upperTranslation - translation;
			float bias = 0.0f, massScale = 1.0f, impulseScale = 0.0f;
			if ( C > 0.0f )
				bias = C * W.inv_h;
			else if ( useBias )
			{
				bias = jointSoftness.biasRate * C;
				massScale = jointSoftness.massScale;
				impulseScale = jointSoftness.impulseScale;
			}
			float oldImpulse = joint.upperImpulse;
			float Cdot = b2Dot( axisA, b2Sub( vA, vB ) ) + a1 * wA - a2 * wB;
			float impulse = -joint.axialMass * massScale * ( Cdot + bias ) - impulseScale * oldImpulse;
			joint.upperImpulse = b2MaxFloat( oldImpulse + impulse, 0.0f );
			impulse = joint.upperImpulse - oldImpulse;
			b2Vec2 P = b2MulSV( impulse, axisA );
			float LA = impulse * a1, LB = impulse * a2;
			vA = b2MulAdd( vA, mA, P );
			wA += iA * LA;
			vB = b2MulSub( vB, mB, P );
			wB -= iB * LB;
		}
	}
	// the perpendicular and angular constraint, solved as a block
	{
		b2Vec2 perpA = b2LeftPerp( axisA );
		float s1 = b2Cross( b2Add( d, rA ), perpA );
		float s2 = b2Cross( rB, perpA );
		b2Vec2 Cdot;
		Cdot.x = b2Dot( perpA, b2Sub( vB, vA ) ) + s2 * wB - s1 * wA;
		Cdot.y = wB - wA;
		b2Vec2 bias = { 0.0f, 0.0f };
		float massScale = 1.0f, impulseScale = 0.0f;
		if ( useBias )
		{
			b2Vec2 C;
			C.x = b2Dot( perpA, d );
			C.y = relative_angle( body_dq( B ), body_dq( A ) ) + joint.deltaAngle;
			bias = b2MulSV( jointSoftness.biasRate, C );
			massScale = jointSoftness.massScale;
			impulseScale = jointSoftness.impulseScale;
		}
		float k11 = mA + mB + iA * s1 * s1 + iB * s2 * s2;
		float k12 = iA * s1 + iB * s2;
		float k22 = iA + iB;
		if ( k22 == 0.0f )
			k22 = 1.0f; // for bodies with fixed rotation
		b2Mat22 K;
		K.cx = b2Vec2{ k11, k12 };
		K.cy = b2Vec2{ k12, k22 };
		b2Vec2 b = solve22( K, b2Add( Cdot, bias ) );
		b2Vec2 impulse;
		impulse.x = -massScale * b.x - impulseScale * joint.impulse.x;
		impulse.y = -massScale * b.y - impulseScale * joint.impulse.y;
		joint.impulse.x += impulse.x;
		joint.impulse.y += impulse.y;
		b2Vec2 P = b2MulSV( impulse.x, perpA );
		float LA = impulse.x * s1 + impulse.y;
		float LB = impulse.x * s2 + impulse.y;
		apply_axis( vA, wA, vB, wB, mA, iA, mB, iB, P, LA, LB );
	}
	store_bodies( A, B, vA, wA, vB, wB );
}

// ---- wheel (src/wheel_joint.c:183-464) ------------------------------------------------------

__device__ inline void prepare_wheel( DevJoint& base, const DevBodySim& simA, const DevBodySim& simB, const DevWorld& W )
{
	DevWheelJoint& joint = base.wheel;
	float mA = simA.invMass, iA = simA.invI, mB = simB.invMass, iB = simB.invI;
	b2Rot qA = sim_rot( simA ), qB = sim_rot( simB );
	joint.anchorA = b2RotateVector( qA, b2Sub( base.localOriginAnchorA, sim_local_center( simA ) ) );
	joint.anchorB = b2RotateVector( qB, b2Sub( base.localOriginAnchorB, sim_local_center( simB ) ) );
	joint.axisA = b2RotateVector( qA, joint.localAxisA );
	joint.deltaCenter = b2Sub( sim_center( simB ), sim_center( simA ) );
	b2Vec2 rA = joint.anchorA, rB = joint.anchorB;
	b2Vec2 d = b2Add( joint.deltaCenter, b2Sub( rB, rA ) );
	b2Vec2 axisA = joint.axisA;
	b2Vec2 perpA = b2LeftPerp( axisA );
	float s1 = b2Cross( b2Add( d, rA ), perpA );
	float s2 = b2Cross( rB, perpA );
	float kp = mA + mB + iA * s1 * s1 + iB * s2 * s2;
	joint.perpMass = kp > 0.0f ? 1.0f / kp : 0.0f;
	float a1 = b2Cross( b2Add( d, rA ), axisA );
	float a2 = b2Cross( rB, axisA );
	float ka = mA + mB + iA * a1 * a1 + iB * a2 * a2;
	joint.axialMass = ka > 0.0f ? 1.0f / ka : 0.0f;
	joint.springSoftness = make_soft( joint.hertz, joint.dampingRatio, W.h );
	float km = iA + iB;
	joint.motorMass = km > 0.0f ? 1.0f / km : 0.0f;
	if ( ( W.flags & kWorldWarmStart ) == 0 )
	{
		joint.perpImpulse = 0.0f;
		joint.springImpulse = 0.0f;
		joint.motorImpulse = 0.0f;
		joint.lowerImpulse = 0.0f;
		joint.upperImpulse = 0.0f;
	}
}

__device__ inline void warm_start_wheel( const DevJoint& base, Body& A, Body& B )
{
	const DevWheelJoint& joint = base.wheel;
	float mA = base.invMassA, mB = base.invMassB, iA = base.invIA, iB = base.invIB;
	b2Vec2 rA = b2RotateVector( body_dq( A ), joint.anchorA );
	b2Vec2 rB = b2RotateVector( body_dq( B ), joint.anchorB );
	b2Vec2 d = axis_separation( A, B, joint.deltaCenter, rA, rB );
	b2Vec2 axisA = b2RotateVector( body_dq( A ), joint.axisA );
	b2Vec2 perpA = b2LeftPerp( axisA );
	float a1 = b2Cross( b2Add( d, rA ), axisA );
	float a2 = b2Cross( rB, axisA );
	float s1 = b2Cross( b2Add( d, rA ), perpA );
	float s2 = b2Cross( rB, perpA );
	float axialImpulse = joint.springImpulse + joint.lowerImpulse - joint.upperImpulse;
	b2Vec2 P = b2Add( b2MulSV( axialImpulse, axisA ), b2MulSV( joint.perpImpulse, perpA ) );
	float LA = axialImpulse * a1 + joint.perpImpulse * s1 + joint.motorImpulse;
	float LB = axialImpulse * a2 + joint.perpImpulse * s2 + joint.motorImpulse;
	A.vx = A.vx - mA * P.x;
	A.vy = A.vy - mA * P.y;
	A.w -= iA * LA;
	B.vx = B.vx + mB * P.x;
	B.vy = B.vy + mB * P.y;
	B.w += iB * LB;
}

__device__ inline void solve_wheel( DevJoint& base, Body& A, Body& B, const DevWorld& W, bool useBias )
{
	DevWheelJoint& joint = base.wheel;
	float mA = base.invMassA, mB = base.invMassB, iA = base.invIA, iB = base.invIB;
	b2Vec2 vA = { A.vx, A.vy }, vB = { B.vx, B.vy };
	float wA = A.w, wB = B.w;
	const DevSoft jointSoftness = joint_softness( W );
	const bool fixedRotation = ( iA + iB == 0.0f );
	b2Vec2 rA = b2RotateVector( body_dq( A ), joint.anchorA );
	b2Vec2 rB = b2RotateVector( body_dq( B ), joint.anchorB );
	b2Vec2 d = axis_separation( A, B, joint.deltaCenter, rA, rB );
	b2Vec2 axisA = b2RotateVector( body_dq( A ), joint.axisA );
	float translation = b2Dot( axisA, d );
	float a1 = b2Cross( b2Add( d, rA ), axisA );
	float a2 = b2Cross( rB, axisA );
	if ( ( base.flags & kJointMotor ) && fixedRotation == false )
	{
		float Cdot = wB - wA - joint.motorSpeed;
		float impulse = -joint.motorMass * Cdot;
		float oldImpulse = joint.motorImpulse;
		float maxImpulse = W.h * joint.maxMotorTorque;
		joint.motorImpulse = b2ClampFloat( joint.motorImpulse + impulse, -maxImpulse, maxImpulse );
		impulse = joint.motorImpulse - oldImpulse;
		wA -= iA * impulse;
		wB += iB * impulse;
	}
	if ( base.flags & kJointSpring )
	{
		float C = translation;
		float bias = joint.springSoftness.biasRate * C;
		float massScale = joint.springSoftness.massScale;
		float impulseScale = joint.springSoftness.impulseScale;
		float Cdot = b2Dot( axisA, b2Sub( vB, vA ) ) + a2 * wB - a1 * wA;
		float impulse = -massScale * joint.axialMass * ( Cdot + bias ) - impulseScale * joint.springImpulse;
		joint.springImpulse += impulse;
		apply_axis( vA, wA, vB, wB, mA, iA, mB, iB, b2MulSV( impulse, axisA ), impulse * a1, impulse * a2 );
	}
	if ( base.flags & kJointLimit )
	{
		// lower
		{
			float C = translation - joint.lowerTranslation;
			float bias = 0.0f, massScale = 1.0f, impulseScale = 0.0f;
			if ( C > 0.0f )
				bias = C * W.inv_h;
			else if ( useBias )
			{
				bias = jointSoftness.biasRate * C;
				massScale = jointSoftness.massScale;
				impulseScale = jointSoftness.impulseScale;
			}
			float Cdot = b2Dot( axisA, b2Sub( vB, vA ) ) + a2 * wB - a1 * wA;
			float impulse = -massScale * joint.axialMass * ( Cdot + bias ) - impulseScale * joint.lowerImpulse;
			float oldImpulse = joint.lowerImpulse;
			joint.lowerImpulse = b2MaxFloat( oldImpulse + impulse, 0.0f );
			impulse = joint.lowerImpulse - oldImpulse;
			apply_axis( vA, wA, vB, wB, mA, iA, mB, iB, b2MulSV( impulse, axisA ), impulse * a1, impulse * a2 );
		}
		// upper (signs flipped)
		{
			float C = joint.upperTranslation - translation;
			float bias = 0.0f, massScale = 1.0f, impulseScale = 0.0f;
			if ( C > 0.0f )
				bias = C * W.inv_h;
			else if ( useBias )
			{
				bias = jointSoftness.biasRate * C;
				massScale = jointSoftness.massScale;
				impulseScale = jointSoftness.impulseScale;
			}
			float Cdot = b2Dot( axisA, b2Sub( vA, vB ) ) + a1 * wA - a2 * wB;
			float impulse = -massScale * joint.axialMass * ( Cdot + bias ) - impulseScale * joint.upperImpulse;
			float oldImpulse = joint.upperImpulse;
			joint.upperImpulse = b2MaxFloat( oldImpulse + impulse, 0.0f );
			impulse = joint.upperImpulse - oldImpulse;
			b2Vec2 P = b2MulSV( impulse, axisA );
			float LA = impulse * a1, LB = impulse * a2;
			vA = b2MulAdd( vA, mA, P );
			wA += iA * LA;
			vB = b2MulSub( vB, mB, P );
			wB -= iB * LB;
		}
	}
	// point to line constraint
	{
		b2Vec2 perpA = b2LeftPerp( axisA );
		float bias = 0.0f, massScale = 1.0f, impulseScale = 0.0f;
		if ( useBias )
		{
			float C = b2Dot( perpA, d );
			bias = jointSoftness.biasRate * C;
			massScale = jointSoftness.massScale;
			impulseScale = jointSoftness.impulseScale;
		}
		float s1 = b2Cross( b2Add( d, rA ), perpA );
		float s2 = b2Cross( rB, perpA );
		float Cdot = b2Dot( perpA, b2Sub( vB, vA ) ) + s2 * wB - s1 * wA;
		float impulse = -massScale * joint.perpMass * ( Cdot + bias ) - impulseScale * joint.perpImpulse;
		joint.perpImpulse += impulse;
		apply_axis( vA, wA, vB, wB, mA, iA, mB, iB, b2MulSV( impulse, perpA ), impulse * s1, impulse * s2 );
	}
	store_bodies( A, B, vA, wA, vB, wB );
}

// ---- motor (src/motor_joint.c:100-270) ------------------------------------------------------

__device__ inline void prepare_motor( DevJoint& base, const DevBodySim& simA, const DevBodySim& simB, const DevWorld& W )
{
	DevMotorJoint& joint = base.motor;
	float mA = simA.invMass, iA = simA.invI, mB = simB.invMass, iB = simB.invI;
	joint.anchorA = b2RotateVector( sim_rot( simA ), b2Sub( base.localOriginAnchorA, sim_local_center( simA ) ) );
	joint.anchorB = b2RotateVector( sim_rot( simB ), b2Sub( base.localOriginAnchorB, sim_local_center( simB ) ) );
	joint.deltaCenter = b2Sub( b2Sub( sim_center( simB ), sim_center( simA ) ), joint.linearOffset );
	joint.deltaAngle = relative_angle( sim_rot( simB ), sim_rot( simA ) ) - joint.angularOffset;
	joint.deltaAngle = unwind_angle( joint.deltaAngle );
	b2Vec2 rA = joint.anchorA, rB = joint.anchorB;
	b2Mat22 K;
	K.cx.x = mA + mB + rA.y * rA.y * iA + rB.y * rB.y * iB;
	K.cx.y = -rA.y * rA.x * iA - rB.y * rB.x * iB;
	K.cy.x = K.cx.y;
	K.cy.y = mA + mB + rA.x * rA.x * iA + rB.x * rB.x * iB;
	joint.linearMass = inverse22( K );
	float ka = iA + iB;
	joint.angularMass = ka > 0.0f ? 1.0f / ka : 0.0f;
	if ( ( W.flags & kWorldWarmStart ) == 0 )
	{
		joint.linearImpulse = b2Vec2{ 0.0f, 0.0f };
		joint.angularImpulse = 0.0f;
	}
}

__device__ inline void warm_start_motor( const DevJoint& base, Body& A, Body& B )
{
	const DevMotorJoint& joint = base.motor;
	float mA = base.invMassA, mB = base.invMassB, iA = base.invIA, iB = base.invIB;
	b2Vec2 rA = b2RotateVector( body_dq( A ), joint.anchorA );
	b2Vec2 rB = b2RotateVector( body_dq( B ), joint.anchorB );
	A.vx = A.vx - mA * joint.linearImpulse.x;
	A.vy = A.vy - mA * joint.linearImpulse.y;
	A.w -= iA * ( b2Cross( rA, joint.linearImpulse ) + joint.angularImpulse );
	B.vx = B.vx + mB * joint.linearImpulse.x;
	B.vy = B.vy + mB * joint.linearImpulse.y;
	B.w += iB * ( b2Cross( rB, joint.linearImpulse ) + joint.angularImpulse );
}

__device__ inline void solve_motor( DevJoint& base, Body& A, Body& B, const DevWorld& W )
{
	DevMotorJoint& joint = base.motor;
	float mA = base.invMassA, mB = base.invMassB, iA = base.invIA, iB = base.invIB;
	b2Vec2 vA = { A.vx, A.vy }, vB = { B.vx, B.vy };
	float wA = A.w, wB = B.w;
	// angular constraint
	{
		float angularSeperation = relative_angle( body_dq( B ), body_dq( A ) ) + joint.deltaAngle;
		angularSeperation = unwind_angle( angularSeperation );
		float angularBias = W.inv_h * joint.correctionFactor * angularSeperation;
		float Cdot = wB - wA;
		float impulse = -joint.angularMass * ( Cdot + angularBias );
		float oldImpulse = joint.angularImpulse;
		float maxImpulse = W.h * joint.maxTorque;
		joint.angularImpulse = b2ClampFloat( joint.angularImpulse + impulse, -maxImpulse, maxImpulse );
		impulse = joint.angularImpulse - oldImpulse;
		wA -= iA * impulse;
		wB += iB * impulse;
	}
	// linear constraint
	{
		b2Vec2 rA = b2RotateVector( body_dq( A ), joint.anchorA );
		b2Vec2 rB = b2RotateVector( body_dq( B ), joint.anchorB );
		b2Vec2 ds = b2Add( b2Sub( b2Vec2{ B.dpx, B.dpy }, b2Vec2{ A.dpx, A.dpy } ), b2Sub( rB, rA ) );
		b2Vec2 linearSeparation = b2Add( joint.deltaCenter, ds );
		b2Vec2 linearBias = b2MulSV( W.inv_h * joint.correctionFactor, linearSeparation );
		b2Vec2 Cdot = b2Sub( b2Add( vB, b2CrossSV( wB, rB ) ), b2Add( vA, b2CrossSV( wA, rA ) ) );
		b2Vec2 b = mul_mv( joint.linearMass, b2Add( Cdot, linearBias ) );
		b2Vec2 impulse = { -b.x, -b.y };
		b2Vec2 oldImpulse = joint.linearImpulse;
		float maxImpulse = W.h * joint.maxForce;
		joint.linearImpulse = b2Add( joint.linearImpulse, impulse );
		if ( b2LengthSquared( joint.linearImpulse ) > maxImpulse * maxImpulse )
		{
			joint.linearImpulse = normalize( joint.linearImpulse );
			joint.linearImpulse.x *= maxImpulse;
			joint.linearImpulse.y *= maxImpulse;
		}
		impulse = b2Sub( joint.linearImpulse, oldImpulse );
		vA = b2MulSub( vA, mA, impulse );
		wA -= iA * b2Cross( rA, impulse );
		vB = b2MulAdd( vB, mB, impulse );
		wB += iB * b2Cross( rB, impulse );
	}
	store_bodies( A, B, vA, wA, vB, wB );
}

// ---- mouse (src/mouse_joint.c:77-230): body B only ------------------------------------------

__device__ inline void prepare_mouse( DevJoint& base, const DevBodySim& simB, const DevWorld& W )
{
	DevMouseJoint& joint = base.mouse;
	joint.anchorB = b2RotateVector( sim_rot( simB ), b2Sub( base.localOriginAnchorB, sim_local_center( simB ) ) );
	joint.linearSoftness = make_soft( joint.hertz, joint.dampingRatio, W.h );
	float angularHertz = 0.5f;
	float angularDampingRatio = 0.1f;
	joint.angularSoftness = make_soft( angularHertz, angularDampingRatio, W.h );
	b2Vec2 rB = joint.anchorB;
	float mB = simB.invMass, iB = simB.invI;
	b2Mat22 K;
	K.cx.x = mB + iB * rB.y * rB.y;
	K.cx.y = -iB * rB.x * rB.y;
	K.cy.x = K.cx.y;
	K.cy.y = mB + iB * rB.x * rB.x;
	joint.linearMass = inverse22( K );
	joint.deltaCenter = b2Sub( sim_center( simB ), joint.targetA );
	if ( ( W.flags & kWorldWarmStart ) == 0 )
	{
		joint.linearImpulse = b2Vec2{ 0.0f, 0.0f };
		joint.angularImpulse = 0.0f;
	}
}

__device__ inline void warm_start_mouse( const DevJoint& base, Body& B )
{
	const DevMouseJoint& joint = base.mouse;
	float mB = base.invMassB, iB = base.invIB;
	b2Vec2 vB = { B.vx, B.vy };
	float wB = B.w;
	b2Vec2 rB = b2RotateVector( body_dq( B ), joint.anchorB );
	vB = b2MulAdd( vB, mB, joint.linearImpulse );
	wB += iB * ( b2Cross( rB, joint.linearImpulse ) + joint.angularImpulse );
	B.vx = vB.x;
	B.vy = vB.y;
	B.w = wB;
}

__device__ inline void solve_mouse( DevJoint& base, Body& B, const DevWorld& W )
{
	DevMouseJoint& joint = base.mouse;
	float mB = base.invMassB, iB = base.invIB;
	b2Vec2 vB = { B.vx, B.vy };
	float wB = B.w;
	// soft angular constraint
	{
		float massScale = joint.angularSoftness.massScale;
		float impulseScale = joint.angularSoftness.impulseScale;
		float impulse = iB > 0.0f ? -wB / iB : 0.0f;
		impulse = massScale * impulse - impulseScale * joint.angularImpulse;
		joint.angularImpulse += impulse;
		wB += iB * impulse;
	}
	float maxImpulse = joint.maxForce * W.h;
	{
		b2Vec2 rB = b2RotateVector( body_dq( B ), joint.anchorB );
		b2Vec2 Cdot = b2Add( vB, b2CrossSV( wB, rB ) );
		b2Vec2 separation = b2Add( b2Add( b2Vec2{ B.dpx, B.dpy }, rB ), joint.deltaCenter );
		b2Vec2 bias = b2MulSV( joint.linearSoftness.biasRate, separation );
		float massScale = joint.linearSoftness.massScale;
		float impulseScale = joint.linearSoftness.impulseScale;
		b2Vec2 b = mul_mv( joint.linearMass, b2Add( Cdot, bias ) );
		b2Vec2 impulse;
		impulse.x = -massScale * b.x - impulseScale * joint.linearImpulse.x;
		impulse.y = -massScale * b.y - impulseScale * joint.linearImpulse.y;
		b2Vec2 oldImpulse = joint.linearImpulse;
		joint.linearImpulse.x += impulse.x;
		joint.linearImpulse.y += impulse.y;
		float mag = b2Length( joint.linearImpulse );
		if ( mag > maxImpulse )
			joint.linearImpulse = b2MulSV( maxImpulse, normalize( joint.linearImpulse ) );
		impulse.x = joint.linearImpulse.x - oldImpulse.x;
		impulse.y = joint.linearImpulse.y - oldImpulse.y;
		vB = b2MulAdd( vB, mB, impulse );
		wB += iB * b2Cross( rB, impulse );
	}
	B.vx = vB.x;
	B.vy = vB.y;
	B.w = wB;
}

// ---- dispatch (src/joint.c:983-1100) -------------------------------------------------------

// prepare: masses and the per-step terms from the sims (src/revolute_joint.c:201-269 and its like)
__device__ inline void prepare_joint( DevJoint& base, const DevBodySim& simA, const DevBodySim& simB, const DevWorld& W )
{
	base.invMassA = simA.invMass;
	base.invMassB = simB.invMass;
	base.invIA = simA.invI;
	base.invIB = simB.invI;
	switch ( base.type )
	{
		case b2_distanceJoint:
			prepare_distance( base, simA, simB, W );
			break;
		case b2_motorJoint:
			prepare_motor( base, simA, simB, W );
			break;
		case b2_mouseJoint:
			prepare_mouse( base, simB, W );
			break;
		case b2_prismaticJoint:
			prepare_prismatic( base, simA, simB, W );
			break;
		case b2_revoluteJoint:
			prepare_revolute( base, simA, simB, W );
			break;
		case b2_weldJoint:
			prepare_weld( base, simA, simB, W );
			break;
		case b2_wheelJoint:
			prepare_wheel( base, simA, simB, W );
			break;
		default:
			break;
	}
}

template <class S>
__device__ inline void warm_start_joint( const S& C, const DevJoint& base, int ia, int ib )
{
	Body A = load_body( C, ia ), B = load_body( C, ib );
	switch ( base.type )
	{
		case b2_distanceJoint:
			warm_start_distance( base, A, B );
			break;
		case b2_motorJoint:
			warm_start_motor( base, A, B );
			break;
		case b2_mouseJoint:
			warm_start_mouse( base, B );
			break;
		case b2_prismaticJoint:
			warm_start_prismatic( base, A, B );
			break;
		case b2_revoluteJoint:
			warm_start_revolute( base, A, B );
			break;
		case b2_weldJoint:
			warm_start_weld( base, A, B );
			break;
		case b2_wheelJoint:
			warm_start_wheel( base, A, B );
			break;
		default:
			break;
	}
	store_velocity( C, ia, A );
	store_velocity( C, ib, B );
}

template <class S>
__device__ inline void solve_joint( const S& C, DevJoint& base, int ia, int ib, const DevWorld& W, bool useBias )
{
	Body A = load_body( C, ia ), B = load_body( C, ib );
	switch ( base.type )
	{
		case b2_distanceJoint:
			solve_distance( base, A, B, W, useBias );
			break;
		case b2_motorJoint:
			solve_motor( base, A, B, W );
			break;
		case b2_mouseJoint:
			solve_mouse( base, B, W );
			break;
		case b2_prismaticJoint:
			solve_prismatic( base, A, B, W, useBias );
			break;
		case b2_revoluteJoint:
			solve_revolute( base, A, B, W, useBias );
			break;
		case b2_weldJoint:
			solve_weld( base, A, B, useBias );
			break;
		case b2_wheelJoint:
			solve_wheel( base, A, B, W, useBias );
			break;
		default:
			break;
	}
	store_velocity( C, ia, A );
	store_velocity( C, ib, B );
}

} // namespace b2mi
