// Replay and adjoint of the solver step, for gradients of a recorded run (CDNA4, gfx950).
//
// A taped run records, for every step, the solver's inputs as the step kernel saw them: the
// work items and slot lists, and a snapshot of every slot's contact (the narrowphase
// manifold, held fixed within the step, with the warm-start impulses it inherited and the
// index of the previous point each came from). Body poses and velocities are checkpointed every
// K steps. The backward pass walks the steps in reverse:
//
//  k_replay_step   re-runs the solver of one step from its start pose (the same device
//                  functions as the step kernel, on the snapshot instead of the live contacts),
//                  producing the next step's start pose; used to rebuild the states between
//                  checkpoints;
//  k_adjoint_step  re-runs one step while taping every constraint application (the two
//                  bodies' states and the four impulses before it), then sweeps the tape in
//                  reverse: finalize, restitution, and per substep (last first) relax,
//                  integrate positions, solve, warm start, integrate velocities, and finally
//                  prepare. Each stage's vector-Jacobian product follows the forward
//                  arithmetic of solver.hpp branch for branch (clamps, the speculative /
//                  soft bias switch, the friction cone, the restitution threshold).
//
// Gradients flow to body centers, rotations, linear and angular velocities at every step
// boundary, to the impulses carried from one step's store to the next step's warm start, and
// to each contact's friction and restitution coefficients (summed per world). The manifold
// (normal, anchors, separation) is a constant of each step. Items with overflow constraints
// are not supported by the adjoint (the host refuses to tape them).
#include <hip/hip_runtime.h>

#include "b2mi_layout.h"
#include "solver.hpp"

#include <type_traits>

namespace b2mi
{

// ---- per-item scratch layout of the adjoint kernel (floats) ----
//  tape of constraint applications: (3 S + 1) passes x slots x kApplFloats
//  tape of bodies: 2 S records x bodies x 8 (before integrate velocities, before positions)
//  body adjoints: bodies x 12 (v.x, v.y, w, -, dp.x, dp.y, dq.c, dq.s, c.x, c.y, q.c, q.s)
//  slot adjoints: slots x 8 (lambda_n0, lambda_n1, lambda_t0, lambda_t1, relVel0, relVel1, -, -)
constexpr int kApplFloats = 20;

__host__ __device__ inline int adjointScratchFloats( int bodyCount, int slotCount, int subSteps )
{
	return kApplFloats * slotCount * ( 3 * subSteps + 1 ) + 16 * subSteps * bodyCount + 12 * bodyCount + 8 * slotCount;
}

struct AdjScratch
{
	float* appl;   // tape of applications
	float* bodyT;  // tape of bodies
	float* bAdj;   // body adjoints
	float* sAdj;   // slot adjoints
	int slots, bodies;

	// Every record starts on a 16-byte boundary (the scratch base is 256-byte aligned and every
	// region and record is a multiple of 4 floats), so neighbouring fields move as one vector.
	__device__ inline float* applAt( int pass, int k ) const
	{
		return static_cast<float*>( __builtin_assume_aligned( appl + ( (size_t)pass * slots + k ) * kApplFloats, 16 ) );
	}
	__device__ inline float* bodyAt( int record, int i ) const
	{
		return static_cast<float*>( __builtin_assume_aligned( bodyT + ( (size_t)record * bodies + i ) * 8, 16 ) );
	}
	__device__ inline float* bodyAdj( int i ) const
	{
		return static_cast<float*>( __builtin_assume_aligned( bAdj + (size_t)12 * i, 16 ) );
	}
	__device__ inline float* slotAdj( int k ) const
	{
		return static_cast<float*>( __builtin_assume_aligned( sAdj + (size_t)8 * k, 16 ) );
	}
};

__device__ inline AdjScratch carve( float* base, int bodies, int slots, int subSteps )
{
	AdjScratch a;
	a.slots = slots;
	a.bodies = bodies;
	a.appl = base;
	a.bodyT = a.appl + (size_t)kApplFloats * slots * ( 3 * subSteps + 1 );
	a.bAdj = a.bodyT + (size_t)16 * subSteps * bodies;
	a.sAdj = a.bAdj + (size_t)12 * bodies;
	return a;
}

// ---- staging of poses ----

// body state of the step kernel (v, w, flags | dp, dq) from a step-start pose
template <class S>
__device__ inline void stage_pose( const S& C, int i, const StepPose& p )
{
	C.b0( i ) = make_float4( p.vx, p.vy, p.w, 0.0f );
	C.b1( i ) = make_float4( 0.0f, 0.0f, 1.0f, 0.0f );
}

// the pose part of finalize (src/solver.c:165-204): center += dp, q = normalize(dq * q)
__device__ inline StepPose finalize_pose( const StepPose& p, float4 s0, float4 s1 )
{
	StepPose out;
	b2Vec2 center = b2Add( v2( p.cx, p.cy ), v2( s1.x, s1.y ) );
	b2Rot dq;
	dq.c = s1.z;
	dq.s = s1.w;
	b2Rot q0;
	q0.c = p.qc;
	q0.s = p.qs;
	b2Rot q = b2NormalizeRot( b2MulRot( dq, q0 ) );
	out.cx = center.x;
	out.cy = center.y;
	out.qc = q.c;
	out.qs = q.s;
	out.vx = s0.x;
	out.vy = s0.y;
	out.w = s0.z;
	out.pad = 0.0f;
	return out;
}

// ---- one step of the solver from a pose, optionally taping ----

struct NoTape
{
	__device__ inline void appl( int, int, const Body&, const Body&, const float* ) const
	{
	}
	__device__ inline void body( int, int, float4, float4 ) const
	{
	}
	static constexpr bool on = false;
};

struct Tape
{
	AdjScratch a;
	__device__ inline void appl( int pass, int k, const Body& bA, const Body& bB, const float* lam ) const
	{
		float* t = a.applAt( pass, k );
		t[0] = bA.vx;
		t[1] = bA.vy;
		t[2] = bA.w;
		t[3] = bA.dpx;
		t[4] = bA.dpy;
		t[5] = bA.dqc;
		t[6] = bA.dqs;
		t[7] = bB.vx;
		t[8] = bB.vy;
		t[9] = bB.w;
		t[10] = bB.dpx;
		t[11] = bB.dpy;
		t[12] = bB.dqc;
		t[13] = bB.dqs;
		t[14] = lam[0];
		t[15] = lam[1];
		t[16] = lam[2];
		t[17] = lam[3];
	}
	__device__ inline void body( int record, int i, float4 s0, float4 s1 ) const
	{
		float* t = a.bodyAt( record, i );
		t[0] = s0.x;
		t[1] = s0.y;
		t[2] = s0.z;
		t[3] = s0.w;
		t[4] = s1.x;
		t[5] = s1.y;
		t[6] = s1.z;
		t[7] = s1.w;
	}
	static constexpr bool on = true;
};

template <class S, class T>
__device__ inline void tape_appl( const S& C, const T& tape, int pass, int k, int la, int lb )
{
	if ( T::on )
	{
		float lam[4] = { C.at( PF( NormalImpulse, 0 ), k ), C.at( PF( NormalImpulse, 1 ), k ), C.at( PF( TangentImpulse, 0 ), k ),
						 C.at( PF( TangentImpulse, 1 ), k ) };
		tape.appl( pass, k, load_body( C, la ), load_body( C, lb ), lam );
	}
}

// The solver of one step for one work item without overflow constraints, from the snapshot
// contacts `snap` (slot order) and the start poses; the same stage sequence and device
// functions as the step kernel's solve_item.
template <class S, class T>
__device__ void replay_item( const S& C, const T& tape, const DevWorld& W, const WorkItem& wi, const int* __restrict__ bl,
							 const DevContact* __restrict__ snap, const StepPose* __restrict__ poseIn, int simBase )
{
	const int tid = threadIdx.x, nt = blockDim.x;
	const int total = wi.groupStart[kGroups];
	for ( int i = tid; i < wi.bodyCount; i += nt )
	{
		stage_pose( C, i, poseIn[simBase + bl[i]] );
		C.mass( i ) = make_float2( W.sims[bl[i]].invMass, W.sims[bl[i]].invI );
	}
	__syncthreads();
	for ( int k = tid; k < total; k += nt )
	{
		int2 lb = C.bodies( k );
		prepare_slot( C, k, load_prep( snap + k ), lb.x, lb.y, W );
	}
	__syncthreads();

	const float inv_h = W.inv_h, h = W.h;
	const float minBiasVel = -W.pushout;
	const bool warm = ( W.flags & kWorldWarmStart ) != 0;
	const Softness soft = world_softness( W );
	for ( int sub = 0; sub < W.subSteps; ++sub )
	{
		for ( int i = tid; i < wi.bodyCount; i += nt )
		{
			tape.body( 2 * sub, i, C.b0( i ), C.b1( i ) );
			integrate_velocity( C, i, W.sims[bl[i]], W );
		}
		__syncthreads();
		if ( warm )
			for ( int g = 1; g < kGroups; ++g )
			{
				int b = wi.groupStart[g], e = wi.groupStart[g + 1];
				if ( b == e )
					continue;
				for ( int k = b + tid; k < e; k += nt )
				{
					int2 lb = C.bodies( k );
					tape_appl( C, tape, 3 * sub + 0, k, lb.x, lb.y );
					lane_warm_start( C, k, lb.x, lb.y );
				}
				__syncthreads();
			}
		for ( int g = 1; g < kGroups; ++g )
		{
			int b = wi.groupStart[g], e = wi.groupStart[g + 1];
			if ( b == e )
				continue;
			for ( int k = b + tid; k < e; k += nt )
			{
				int2 lb = C.bodies( k );
				tape_appl( C, tape, 3 * sub + 1, k, lb.x, lb.y );
				lane_solve( C, k, lb.x, lb.y, inv_h, minBiasVel, true, soft );
			}
			__syncthreads();
		}
		for ( int i = tid; i < wi.bodyCount; i += nt )
		{
			tape.body( 2 * sub + 1, i, C.b0( i ), C.b1( i ) );
			integrate_position( C, i, h );
		}
		__syncthreads();
		for ( int g = 1; g < kGroups; ++g )
		{
			int b = wi.groupStart[g], e = wi.groupStart[g + 1];
			if ( b == e )
				continue;
			for ( int k = b + tid; k < e; k += nt )
			{
				int2 lb = C.bodies( k );
				tape_appl( C, tape, 3 * sub + 2, k, lb.x, lb.y );
				lane_solve( C, k, lb.x, lb.y, inv_h, minBiasVel, false, soft );
			}
			__syncthreads();
		}
	}
	const float threshold = W.restitutionThreshold;
	for ( int g = 1; g < kGroups; ++g )
	{
		int b = wi.groupStart[g], e = wi.groupStart[g + 1];
		if ( b == e )
			continue;
		for ( int k = b + tid; k < e; k += nt )
		{
			int2 lb = C.bodies( k );
			tape_appl( C, tape, 3 * W.subSteps, k, lb.x, lb.y );
			lane_restitution( C, k, lb.x, lb.y, threshold );
		}
		__syncthreads();
	}
}

// ---- vector-Jacobian products of the stages ----

// adjoint accumulators of one body (a no-op for the static body, index -1)
struct BodyAdj
{
	float* p;
	__device__ inline void addVel( float vx, float vy, float w ) const
	{
		if ( p )
		{
			p[0] += vx;
			p[1] += vy;
			p[2] += w;
		}
	}
	__device__ inline void addPos( float dpx, float dpy, float dqc, float dqs ) const
	{
		if ( p )
		{
			p[4] += dpx;
			p[5] += dpy;
			p[6] += dqc;
			p[7] += dqs;
		}
	}
	__device__ inline float vx() const
	{
		return p ? p[0] : 0.0f;
	}
	__device__ inline float vy() const
	{
		return p ? p[1] : 0.0f;
	}
	__device__ inline float w() const
	{
		return p ? p[2] : 0.0f;
	}
};

__device__ inline BodyAdj body_adj( const AdjScratch& a, int i )
{
	BodyAdj b;
	b.p = i >= 0 ? a.bodyAdj( i ) : nullptr;
	return b;
}

// adjoint of an impulse P = delta * dir applied with lane_apply: d(loss)/d(delta)
__device__ inline float apply_adj( const BodyAdj& A, const BodyAdj& B, float mA, float iA, float mB, float iB, float rAx, float rAy,
								   float rBx, float rBy, float dx, float dy )
{
	// vA -= mA P, wA -= iA (rA x P), vB += mB P, wB += iB (rB x P), with rA x P = rAx Py - rAy Px
	float Px = -mA * A.vx() + mB * B.vx() + iA * rAy * A.w() - iB * rBy * B.w();
	float Py = -mA * A.vy() + mB * B.vy() - iA * rAx * A.w() + iB * rBx * B.w();
	return Px * dx + Py * dy;
}

// adjoint of the relative velocity along dir at the anchors: d = (vB + wB x rB - vA - wA x rA) . dir
__device__ inline void relvel_adj( const BodyAdj& A, const BodyAdj& B, float rAx, float rAy, float rBx, float rBy, float dx, float dy,
								   float g )
{
	float gx = g * dx, gy = g * dy;
	B.addVel( gx, gy, -rBy * gx + rBx * gy );
	A.addVel( -gx, -gy, rAy * gx - rAx * gy );
}

template <class S>
__device__ void lane_warm_start_adj( const S& C, const AdjScratch& a, int k, int la, int lb )
{
	BodyAdj A = body_adj( a, la ), B = body_adj( a, lb );
	float2 MA = body_mass( C, la ), MB = body_mass( C, lb );
	float mA = MA.x, mB = MB.x, iA = MA.y, iB = MB.y;
	float nx = C.at( cfNormalX, k ), ny = C.at( cfNormalY, k );
	float tx = ny, ty = 0.0f - nx;
	float* sa = a.slotAdj( k );
	for ( int j = 0; j < 2; ++j )
	{
		float rAx = C.at( PF( AnchorAX, j ), k ), rAy = C.at( PF( AnchorAY, j ), k );
		float rBx = C.at( PF( AnchorBX, j ), k ), rBy = C.at( PF( AnchorBY, j ), k );
		// velocity adjoints pass through unchanged; P = ni n + ti t
		float Px = -mA * A.vx() + mB * B.vx() + iA * rAy * A.w() - iB * rBy * B.w();
		float Py = -mA * A.vy() + mB * B.vy() - iA * rAx * A.w() + iB * rBx * B.w();
		sa[j] += Px * nx + Py * ny;
		sa[2 + j] += Px * tx + Py * ty;
	}
}

// reverse of lane_solve from its taped inputs; returns d(loss)/d(friction) of this application
template <class S>
__device__ float lane_solve_adj( const S& C, const AdjScratch& a, int pass, int k, int la, int lb, float inv_h, float minBiasVel,
								 bool useBias, const DevWorld& W )
{
	const float* t = a.applAt( pass, k );
	Body bA, bB;
	bA.vx = t[0], bA.vy = t[1], bA.w = t[2], bA.dpx = t[3], bA.dpy = t[4], bA.dqc = t[5], bA.dqs = t[6];
	bB.vx = t[7], bB.vy = t[8], bB.w = t[9], bB.dpx = t[10], bB.dpy = t[11], bB.dqc = t[12], bB.dqs = t[13];
	float ni0[2] = { t[14], t[15] }, ti0[2] = { t[16], t[17] };
	float2 MA = body_mass( C, la ), MB = body_mass( C, lb );
	float mA = MA.x, mB = MB.x, iA = MA.y, iB = MB.y;
	float nx = C.at( cfNormalX, k ), ny = C.at( cfNormalY, k );
	const Softness sw = world_softness( W );
	float biasRate = useBias ? contact_soft( sw, la, lb, 0 ) : 0.0f;
	float massScale = useBias ? contact_soft( sw, la, lb, 1 ) : 1.0f;
	float impulseScale = useBias ? contact_soft( sw, la, lb, 2 ) : 0.0f;
	float friction = C.at( cfFriction, k );
	float tx = ny, ty = 0.0f - nx;
	float rAx[2], rAy[2], rBx[2], rBy[2];
	for ( int j = 0; j < 2; ++j )
	{
		rAx[j] = C.at( PF( AnchorAX, j ), k ), rAy[j] = C.at( PF( AnchorAY, j ), k );
		rBx[j] = C.at( PF( AnchorBX, j ), k ), rBy[j] = C.at( PF( AnchorBY, j ), k );
	}

	// forward again, keeping what the reverse needs (same arithmetic as lane_solve)
	float dpx = bB.dpx - bA.dpx, dpy = bB.dpy - bA.dpy;
	float sN[2], uN[2], nNew[2];
	for ( int j = 0; j < 2; ++j )
	{
		float rsAx = bA.dqc * rAx[j] - bA.dqs * rAy[j], rsAy = bA.dqs * rAx[j] + bA.dqc * rAy[j];
		float rsBx = bB.dqc * rBx[j] - bB.dqs * rBy[j], rsBy = bB.dqs * rBx[j] + bB.dqc * rBy[j];
		float dsx = dpx + ( rsBx - rsAx ), dsy = dpy + ( rsBy - rsAy );
		float s = ( nx * dsx + ny * dsy ) + C.at( PF( BaseSeparation, j ), k );
		float specBias = s * inv_h;
		float softBias = maxr( biasRate * s, minBiasVel );
		float bias = s > 0.0f ? specBias : softBias;
		float dvx = ( bB.vx - bB.w * rBy[j] ) - ( bA.vx - bA.w * rAy[j] );
		float dvy = ( bB.vy + bB.w * rBx[j] ) - ( bA.vy + bA.w * rAx[j] );
		float vn = dvx * nx + dvy * ny;
		float negImpulse = C.at( PF( NormalMass, j ), k ) * ( massScale * ( vn + bias ) ) + impulseScale * ni0[j];
		float u = ni0[j] - negImpulse;
		float newImpulse = maxr( u, 0.0f );
		float impulse = newImpulse - ni0[j];
		sN[j] = s;
		uN[j] = u;
		nNew[j] = newImpulse;
		lane_apply( bA, bB, mA, iA, mB, iB, rAx[j], rAy[j], rBx[j], rBy[j], impulse * nx, impulse * ny );
	}
	float uT[2], maxF[2];
	for ( int j = 0; j < 2; ++j )
	{
		float dvx = ( bB.vx - bB.w * rBy[j] ) - ( bA.vx - bA.w * rAy[j] );
		float dvy = ( bB.vy + bB.w * rBx[j] ) - ( bA.vy + bA.w * rAx[j] );
		float vt = dvx * tx + dvy * ty;
		float negImpulse = C.at( PF( TangentMass, j ), k ) * vt;
		maxF[j] = friction * nNew[j];
		float u = ti0[j] - negImpulse;
		uT[j] = u;
		float newImpulse = maxr( 0.0f - maxF[j], minr( u, maxF[j] ) );
		float impulse = newImpulse - ti0[j];
		lane_apply( bA, bB, mA, iA, mB, iB, rAx[j], rAy[j], rBx[j], rBy[j], impulse * tx, impulse * ty );
	}

	// reverse: friction point 1, 0, then normal point 1, 0
	BodyAdj A = body_adj( a, la ), B = body_adj( a, lb );
	float* sa = a.slotAdj( k );
	float gFriction = 0.0f;
	float nBar[2] = { sa[0], sa[1] }; // adjoint of the normal impulses after this application
	for ( int j = 1; j >= 0; --j )
	{
		float gDelta = apply_adj( A, B, mA, iA, mB, iB, rAx[j], rAy[j], rBx[j], rBy[j], tx, ty );
		float gNew = sa[2 + j] + gDelta; // d/d(new tangent impulse)
		float gOld = -gDelta;			 // d/d(old tangent impulse), through impulse = new - old
		float gU = 0.0f, gMaxF = 0.0f;
		if ( uT[j] > maxF[j] )
			gMaxF = gNew; // clamped to +maxF (min picked maxF)
		else if ( uT[j] < 0.0f - maxF[j] )
			gMaxF = -gNew; // clamped to -maxF
		else
			gU = gNew;
		gFriction += gMaxF * nNew[j];
		nBar[j] += gMaxF * friction;
		gOld += gU;
		float gVt = -C.at( PF( TangentMass, j ), k ) * gU;
		relvel_adj( A, B, rAx[j], rAy[j], rBx[j], rBy[j], tx, ty, gVt );
		sa[2 + j] = gOld;
	}
	float gDp[2] = { 0.0f, 0.0f };
	float gDqA[2] = { 0.0f, 0.0f }, gDqB[2] = { 0.0f, 0.0f };
	for ( int j = 1; j >= 0; --j )
	{
		float gDelta = apply_adj( A, B, mA, iA, mB, iB, rAx[j], rAy[j], rBx[j], rBy[j], nx, ny );
		float gNew = nBar[j] + gDelta;
		float gOld = -gDelta;
		float gU = uN[j] > 0.0f ? gNew : 0.0f;
		gOld += gU;
		float gNeg = -gU;
		float normalMass = C.at( PF( NormalMass, j ), k );
		float gVn = normalMass * massScale * gNeg;
		float gBias = normalMass * massScale * gNeg;
		gOld += impulseScale * gNeg;
		float s = sN[j];
		float gS = 0.0f;
		if ( s > 0.0f )
			gS = inv_h * gBias;
		else if ( biasRate * s > minBiasVel )
			gS = biasRate * gBias;
		// s = n . (dpB - dpA + rot(dqB, rB) - rot(dqA, rA)) + base
		float gdx = gS * nx, gdy = gS * ny;
		gDp[0] += gdx;
		gDp[1] += gdy;
		gDqB[0] += gdx * rBx[j] + gdy * rBy[j];
		gDqB[1] += -gdx * rBy[j] + gdy * rBx[j];
		gDqA[0] -= gdx * rAx[j] + gdy * rAy[j];
		gDqA[1] -= -gdx * rAy[j] + gdy * rAx[j];
		relvel_adj( A, B, rAx[j], rAy[j], rBx[j], rBy[j], nx, ny, gVn );
		sa[j] = gOld;
	}
	B.addPos( gDp[0], gDp[1], gDqB[0], gDqB[1] );
	A.addPos( -gDp[0], -gDp[1], gDqA[0], gDqA[1] );
	return gFriction;
}

// reverse of lane_restitution; returns d(loss)/d(restitution) of this application
template <class S>
__device__ float lane_restitution_adj( const S& C, const AdjScratch& a, int pass, int k, int la, int lb, float threshold )
{
	const float* t = a.applAt( pass, k );
	Body bA, bB;
	bA.vx = t[0], bA.vy = t[1], bA.w = t[2], bA.dpx = t[3], bA.dpy = t[4], bA.dqc = t[5], bA.dqs = t[6];
	bB.vx = t[7], bB.vy = t[8], bB.w = t[9], bB.dpx = t[10], bB.dpy = t[11], bB.dqc = t[12], bB.dqs = t[13];
	float ni0[2] = { t[14], t[15] };
	float2 MA = body_mass( C, la ), MB = body_mass( C, lb );
	float mA = MA.x, mB = MB.x, iA = MA.y, iB = MB.y;
	float nx = C.at( cfNormalX, k ), ny = C.at( cfNormalY, k );
	float restitution = C.at( cfRestitution, k );
	float rAx[2], rAy[2], rBx[2], rBy[2], mass[2], relVel[2], u[2];
	for ( int j = 0; j < 2; ++j )
	{
		rAx[j] = C.at( PF( AnchorAX, j ), k ), rAy[j] = C.at( PF( AnchorAY, j ), k );
		rBx[j] = C.at( PF( AnchorBX, j ), k ), rBy[j] = C.at( PF( AnchorBY, j ), k );
		relVel[j] = C.at( PF( RelativeVelocity, j ), k );
		bool skip = ( relVel[j] + threshold ) > 0.0f || C.at( PF( MaxNormalImpulse, j ), k ) == 0.0f;
		mass[j] = skip ? 0.0f : C.at( PF( NormalMass, j ), k );
		float dvx = ( bB.vx - bB.w * rBy[j] ) - ( bA.vx - bA.w * rAy[j] );
		float dvy = ( bB.vy + bB.w * rBx[j] ) - ( bA.vy + bA.w * rAx[j] );
		float vn = dvx * nx + dvy * ny;
		float negImpulse = mass[j] * ( vn + restitution * relVel[j] );
		u[j] = ni0[j] - negImpulse;
		float newImpulse = maxr( u[j], 0.0f );
		float impulse = newImpulse - ni0[j];
		lane_apply( bA, bB, mA, iA, mB, iB, rAx[j], rAy[j], rBx[j], rBy[j], impulse * nx, impulse * ny );
	}
	BodyAdj A = body_adj( a, la ), B = body_adj( a, lb );
	float* sa = a.slotAdj( k );
	float gRest = 0.0f;
	for ( int j = 1; j >= 0; --j )
	{
		float gDelta = apply_adj( A, B, mA, iA, mB, iB, rAx[j], rAy[j], rBx[j], rBy[j], nx, ny );
		float gNew = sa[j] + gDelta;
		float gOld = -gDelta;
		float gU = u[j] > 0.0f ? gNew : 0.0f;
		gOld += gU;
		float gNeg = -gU;
		float gVn = mass[j] * gNeg;
		gRest += mass[j] * relVel[j] * gNeg;
		sa[4 + j] += mass[j] * restitution * gNeg;
		relvel_adj( A, B, rAx[j], rAy[j], rBx[j], rBy[j], nx, ny, gVn );
		sa[j] = gOld;
	}
	return gRest;
}

// reverse of integrate_velocity for body i from its taped state before the integration
__device__ void integrate_velocity_adj( const AdjScratch& a, int i, const float* before, const DevBodySim& sim, const DevWorld& W )
{
	float* g = a.bodyAdj( i );
	float h = W.h;
	float maxLinearSpeed = W.maxLinearSpeed, maxAngularSpeed = W.maxAngularSpeed;
	float linearDamping = 1.0f / ( 1.0f + h * sim.linearDamping );
	float angularDamping = 1.0f / ( 1.0f + h * sim.angularDamping );
	float mg = sim.mass * sim.gravityScale;
	float fx = sim.force[0] + mg * W.gravityX, fy = sim.force[1] + mg * W.gravityY;
	float hm = h * sim.invMass;
	float vx = hm * fx + linearDamping * before[0];
	float vy = hm * fy + linearDamping * before[1];
	float w = h * sim.invI * sim.torque + angularDamping * before[2];
	float gvx = g[0], gvy = g[1], gw = g[2];
	if ( vx * vx + vy * vy > maxLinearSpeed * maxLinearSpeed )
	{
		// v' = (max / |u|) u
		float len = sqrtf( vx * vx + vy * vy );
		float ux = vx / len, uy = vy / len;
		float d = ux * gvx + uy * gvy;
		float r = maxLinearSpeed / len;
		gvx = r * ( gvx - ux * d );
		gvy = r * ( gvy - uy * d );
	}
	if ( w * w > maxAngularSpeed * maxAngularSpeed )
		gw = 0.0f;
	g[0] = linearDamping * gvx;
	g[1] = linearDamping * gvy;
	g[2] = angularDamping * gw;
}

// reverse of integrate_position for body i from its taped state before the integration
__device__ void integrate_position_adj( const AdjScratch& a, int i, const float* before, float h )
{
	float* g = a.bodyAdj( i );
	float w = before[2], dqc = before[6], dqs = before[7];
	float da = h * w;
	float c2 = dqc - da * dqs;
	float s2 = dqs + da * dqc;
	float mag = sqrtf( s2 * s2 + c2 * c2 );
	float inv = mag > 0.0f ? 1.0f / mag : 0.0f;
	float oc = c2 * inv, os = s2 * inv;
	// dp' = dp + h v
	g[0] += h * g[4];
	g[1] += h * g[5];
	// dq' = normalize(c2, s2)
	float d = oc * g[6] + os * g[7];
	float gc2 = inv * ( g[6] - oc * d );
	float gs2 = inv * ( g[7] - os * d );
	float gda = -dqs * gc2 + dqc * gs2;
	g[6] = gc2 + da * gs2;
	g[7] = -da * gc2 + gs2;
	g[2] += h * gda;
}

// ---- the solver state of a work item: in HBM, or in LDS as the step kernel's LDS store ----

extern __shared__ float4 lds[];

template <class S>
__device__ inline S item_store( const WorkItem& wi, const int2* __restrict__ slotBodies, float* cfBase, float4* priv, float2* mass );

template <>
__device__ inline GlobalStore item_store<GlobalStore>( const WorkItem& wi, const int2* __restrict__ slotBodies, float* cfBase,
													   float4* priv, float2* mass )
{
	GlobalStore C;
	C.rec = reinterpret_cast<float4*>( cfBase );
	C.base = wi.slotStart;
	C.sb = slotBodies + wi.slotStart;
	C.P = priv + 2 * (size_t)wi.bodyStart;
	C.M = mass + wi.bodyStart;
	return C;
}

// Dynamic LDS of 40 (bodyCount + 1) + 16 cfQuads slots bytes: body state and the static body's
// identity state, constraint records, masses. Written here; read after the barrier that ends
// the pose staging of replay_item.
template <>
__device__ inline LdsStore item_store<LdsStore>( const WorkItem& wi, const int2* __restrict__ slotBodies, float*, float4*, float2* )
{
	const int total = wi.groupStart[kGroups];
	LdsStore C;
	C.P = lds;
	C.identity = wi.bodyCount;
	C.rec = lds + 2 * ( wi.bodyCount + 1 );
	C.M = reinterpret_cast<float2*>( C.rec + cfQuads * total );
	for ( int k = threadIdx.x; k < total; k += blockDim.x )
	{
		int2 b = slotBodies[wi.slotStart + k];
		C.rec[k * cfQuads] = make_float4( __builtin_bit_cast( float, b.x ), __builtin_bit_cast( float, b.y ), 0.0f, 0.0f );
	}
	if ( threadIdx.x == 0 )
	{
		C.b0( C.identity ) = make_float4( 0.0f, 0.0f, 0.0f, 0.0f );
		C.b1( C.identity ) = make_float4( 0.0f, 0.0f, 1.0f, 0.0f );
		C.mass( C.identity ) = make_float2( 0.0f, 0.0f );
	}
	return C;
}

// ---- kernels ----

// One workgroup per work item: the step's solver from poseIn to poseOut (no tape).
template <bool kLds>
__global__ void __launch_bounds__( kAdjointMaxBlock ) k_replay_step( const DevWorld* __restrict__ worlds, const WorkItem* __restrict__ items,
																   const int* __restrict__ bodyList, const int2* __restrict__ slotBodies,
																   const DevContact* __restrict__ snap, const StepPose* __restrict__ poseIn,
																   StepPose* __restrict__ poseOut, const int* __restrict__ simBase,
																   float* cfBase, int cfCapacity, float4* priv, float2* mass )
{
	const WorkItem& wi = items[blockIdx.x];
	const DevWorld& W = worlds[wi.world];
	const int* bl = bodyList + wi.bodyStart;
	using Store = std::conditional_t<kLds, LdsStore, GlobalStore>;
	const Store C = item_store<Store>( wi, slotBodies, cfBase, priv, mass );
	NoTape tape;
	replay_item( C, tape, W, wi, bl, snap + wi.slotStart, poseIn, simBase[wi.world] );
	int base = simBase[wi.world];
	for ( int i = threadIdx.x; i < wi.ownedCount; i += blockDim.x )
		poseOut[base + bl[i]] = finalize_pose( poseIn[base + bl[i]], C.b0( i ), C.b1( i ) );
}

// One workgroup per work item: replay with tape, then the reverse sweep. poseIn is the step's
// start pose; adj holds, per body, the adjoint of the step's end pose on entry and of its start
// pose on exit (cx, cy, qc, qs, vx, vy, w). lamOut holds, per slot, the adjoints of the stored
// impulses (normal 0, 1, tangent 0, 1); lamPrev receives the adjoints of the previous step's
// stored impulses through the warm-start links (prevSlot: per slot of this step, the previous
// step's slot of the same contact, or -1). paramGrad accumulates per world d/d(friction) and
// d/d(restitution).
template <bool kAdjLds>
__global__ void __launch_bounds__( kAdjointMaxBlock ) k_adjoint_step( const DevWorld* __restrict__ worlds, const WorkItem* __restrict__ items,
																	const int* __restrict__ bodyList, const int2* __restrict__ slotBodies,
																	const DevContact* __restrict__ snap, const StepPose* __restrict__ poseIn,
																	StepPose* __restrict__ adj, const int* __restrict__ simBase,
																	const float4* __restrict__ lamOut, float* __restrict__ lamPrev,
																	const int* __restrict__ prevSlot, float* __restrict__ paramGrad,
																	float* cfBase, int cfCapacity, float4* priv, float2* mass, float* scratch,
																	const int* __restrict__ scratchOffset )
{
	const WorkItem& wi = items[blockIdx.x];
	const DevWorld& W = worlds[wi.world];
	const int* bl = bodyList + wi.bodyStart;
	const int tid = threadIdx.x, nt = blockDim.x;
	const int total = wi.groupStart[kGroups];
	const int base = simBase[wi.world];
	const int S = W.subSteps;
	// with kAdjLds the solver state is in LDS too
	using Store = std::conditional_t<kAdjLds, LdsStore, GlobalStore>;
	const Store C = item_store<Store>( wi, slotBodies, cfBase, priv, mass );
	Tape tape;
	tape.a = carve( scratch + scratchOffset[blockIdx.x], wi.bodyCount, total, S );
	if constexpr ( kAdjLds )
	{
		// every item of the launch has 12 bodyCount + 8 slots <= kAdjointLdsFloats (checked by the host)
		__shared__ __align__( 16 ) float adjLds[kAdjointLdsFloats];
		tape.a.bAdj = adjLds;
		tape.a.sAdj = adjLds + (size_t)12 * wi.bodyCount;
	}
	const AdjScratch& a = tape.a;
	const DevContact* sn = snap + wi.slotStart;

	replay_item( C, tape, W, wi, bl, sn, poseIn, base );

	// finalize^T: center' = center + dp, q' = normalize(dq q); velocities pass through
	for ( int i = tid; i < wi.bodyCount; i += nt )
	{
		float* g = a.bodyAdj( i );
		for ( int f = 0; f < 12; ++f )
			g[f] = 0.0f;
		if ( i >= wi.ownedCount )
			continue;
		StepPose ge = adj[base + bl[i]];
		StepPose p0 = poseIn[base + bl[i]];
		float4 s1 = C.b1( i );
		float dqc = s1.z, dqs = s1.w;
		float mc = dqc * p0.qc - dqs * p0.qs;
		float ms = dqs * p0.qc + dqc * p0.qs;
		float mag = sqrtf( ms * ms + mc * mc );
		float inv = mag > 0.0f ? 1.0f / mag : 0.0f;
		float oc = mc * inv, os = ms * inv;
		float d = oc * ge.qc + os * ge.qs;
		float gmc = inv * ( ge.qc - oc * d ), gms = inv * ( ge.qs - os * d );
		g[0] = ge.vx;
		g[1] = ge.vy;
		g[2] = ge.w;
		g[4] = ge.cx;
		g[5] = ge.cy;
		g[6] = gmc * p0.qc + gms * p0.qs;
		g[7] = -gmc * p0.qs + gms * p0.qc;
		g[8] = ge.cx;
		g[9] = ge.cy;
		g[10] = gmc * dqc + gms * dqs;
		g[11] = -gmc * dqs + gms * dqc;
	}
	// store^T
	for ( int k = tid; k < total; k += nt )
	{
		float4 l = lamOut[wi.slotStart + k];
		float* sa = a.slotAdj( k );
		sa[0] = l.x;
		sa[1] = l.y;
		sa[2] = l.z;
		sa[3] = l.w;
		sa[4] = 0.0f;
		sa[5] = 0.0f;
		sa[6] = 0.0f; // d/d(friction) of this slot
		sa[7] = 0.0f; // d/d(restitution) of this slot
	}
	__syncthreads();

	const float inv_h = W.inv_h, h = W.h;
	const float minBiasVel = -W.pushout;
	const bool warm = ( W.flags & kWorldWarmStart ) != 0;
	const float threshold = W.restitutionThreshold;
	for ( int g = kGroups - 1; g >= 1; --g )
	{
		int b = wi.groupStart[g], e = wi.groupStart[g + 1];
		if ( b == e )
			continue;
		for ( int k = b + tid; k < e; k += nt )
		{
			int2 lb = C.bodies( k );
			a.slotAdj( k )[7] += lane_restitution_adj( C, a, 3 * S, k, lb.x, lb.y, threshold );
		}
		__syncthreads();
	}
	for ( int sub = S - 1; sub >= 0; --sub )
	{
		for ( int g = kGroups - 1; g >= 1; --g )
		{
			int b = wi.groupStart[g], e = wi.groupStart[g + 1];
			if ( b == e )
				continue;
			for ( int k = b + tid; k < e; k += nt )
			{
				int2 lb = C.bodies( k );
				a.slotAdj( k )[6] += lane_solve_adj( C, a, 3 * sub + 2, k, lb.x, lb.y, inv_h, minBiasVel, false, W );
			}
			__syncthreads();
		}
		for ( int i = tid; i < wi.bodyCount; i += nt )
			integrate_position_adj( a, i, a.bodyAt( 2 * sub + 1, i ), h );
		__syncthreads();
		for ( int g = kGroups - 1; g >= 1; --g )
		{
			int b = wi.groupStart[g], e = wi.groupStart[g + 1];
			if ( b == e )
				continue;
			for ( int k = b + tid; k < e; k += nt )
			{
				int2 lb = C.bodies( k );
				a.slotAdj( k )[6] += lane_solve_adj( C, a, 3 * sub + 1, k, lb.x, lb.y, inv_h, minBiasVel, true, W );
			}
			__syncthreads();
		}
		if ( warm )
			for ( int g = kGroups - 1; g >= 1; --g )
			{
				int b = wi.groupStart[g], e = wi.groupStart[g + 1];
				if ( b == e )
					continue;
				for ( int k = b + tid; k < e; k += nt )
				{
					int2 lb = C.bodies( k );
					lane_warm_start_adj( C, a, k, lb.x, lb.y );
				}
				__syncthreads();
			}
		for ( int i = tid; i < wi.bodyCount; i += nt )
			integrate_velocity_adj( a, i, a.bodyAt( 2 * sub, i ), W.sims[bl[i]], W );
		__syncthreads();
	}

	// prepare^T: relative velocities from the start velocities; warm-start impulses from the
	// previous step's stored impulses (scaled by warmStartScale), routed through warmFrom
	for ( int k = tid; k < total; k += nt )
	{
		int2 lb = C.bodies( k );
		BodyAdj A = body_adj( a, lb.x ), B = body_adj( a, lb.y );
		float nx = C.at( cfNormalX, k ), ny = C.at( cfNormalY, k );
		float* sa = a.slotAdj( k );
		const DevContact& ct = sn[k];
		int prev = prevSlot[wi.slotStart + k];
		for ( int j = 0; j < 2; ++j )
		{
			if ( j >= ct.manifold.pointCount )
				continue;
			// the relative velocity uses the manifold anchors (equal to the prepared ones)
			float rAx = C.at( PF( AnchorAX, j ), k ), rAy = C.at( PF( AnchorAY, j ), k );
			float rBx = C.at( PF( AnchorBX, j ), k ), rBy = C.at( PF( AnchorBY, j ), k );
			relvel_adj( A, B, rAx, rAy, rBx, rBy, nx, ny, sa[4 + j] );
			uint32_t from = ( ct.warmFrom >> ( 2 * j ) ) & 3u;
			if ( prev >= 0 && from < 2u )
			{
				float* lp = lamPrev + 4 * (size_t)prev;
				lp[from] = W.warmStartScale * sa[j];
				lp[2 + from] = W.warmStartScale * sa[2 + j];
			}
		}
	}
	__syncthreads();

	// start-pose adjoints: center and rotation from finalize, velocities from the sweep
	for ( int i = tid; i < wi.ownedCount; i += nt )
	{
		const float* g = a.bodyAdj( i );
		StepPose o;
		o.cx = g[8];
		o.cy = g[9];
		o.qc = g[10];
		o.qs = g[11];
		o.vx = g[0];
		o.vy = g[1];
		o.w = g[2];
		o.pad = 0.0f;
		adj[base + bl[i]] = o;
	}
	// parameter gradients of this item (fixed-order sum: deterministic)
	if ( tid == 0 )
	{
		float gf = 0.0f, gr = 0.0f;
		for ( int k = 0; k < total; ++k )
		{
			gf += a.slotAdj( k )[6];
			gr += a.slotAdj( k )[7];
		}
		paramGrad[2 * wi.world] += gf;
		paramGrad[2 * wi.world + 1] += gr;
	}
}

// snapshot of the contacts of every slot, in slot order, as the step kernel will read them
__global__ void k_snapshot( const DevWorld* __restrict__ worlds, const WorkItem* __restrict__ items, const int* __restrict__ slotContact,
							DevContact* __restrict__ snap )
{
	const WorkItem& wi = items[blockIdx.x];
	const DevWorld& W = worlds[wi.world];
	const int total = wi.groupStart[kGroups];
	for ( int k = threadIdx.x; k < total; k += blockDim.x )
		snap[wi.slotStart + k] = W.contacts[slotContact[wi.slotStart + k]];
}

// start poses (center, rotation, velocity) of every awake body, per world at simBase
__global__ void k_gather_poses( const DevWorld* __restrict__ worlds, const int* __restrict__ simBase, const int* __restrict__ simCount,
								int worldCount, StepPose* __restrict__ out )
{
	int wIndex = blockIdx.y;
	if ( wIndex >= worldCount )
		return;
	const DevWorld& W = worlds[wIndex];
	const float4* st = reinterpret_cast<const float4*>( W.state );
	for ( int i = blockIdx.x * blockDim.x + threadIdx.x; i < simCount[wIndex]; i += gridDim.x * blockDim.x )
	{
		const DevBodySim& s = W.sims[i];
		float4 v = st[2 * i];
		StepPose p;
		p.cx = s.center[0];
		p.cy = s.center[1];
		p.qc = s.q[0];
		p.qs = s.q[1];
		p.vx = v.x;
		p.vy = v.y;
		p.w = v.z;
		p.pad = 0.0f;
		out[simBase[wIndex] + i] = p;
	}
}

// prevSlot[k] for the slots of step t: the slot the same contact held in step t - 1
__global__ void k_link_slots( const int* __restrict__ slotContact, const int* __restrict__ slotWorld, int slotCount,
							  const int* __restrict__ contactBase, const int2* __restrict__ slotOfContact, int prevStep,
							  int* __restrict__ prevSlot )
{
	int k = blockIdx.x * blockDim.x + threadIdx.x;
	if ( k >= slotCount )
		return;
	// slots between the worlds' regions (world -1) hold no contact
	if ( slotWorld[k] < 0 )
	{
		prevSlot[k] = -1;
		return;
	}
	int2 e = slotOfContact[contactBase[slotWorld[k]] + slotContact[k]];
	prevSlot[k] = e.x == prevStep ? e.y : -1;
}

// slotOfContact[contact] = (step, slot) for the slots of one step
__global__ void k_mark_slots( const int* __restrict__ slotContact, const int* __restrict__ slotWorld, int slotCount,
							  const int* __restrict__ contactBase, int step, int2* __restrict__ slotOfContact )
{
	int k = blockIdx.x * blockDim.x + threadIdx.x;
	if ( k >= slotCount || slotWorld[k] < 0 )
		return;
	slotOfContact[contactBase[slotWorld[k]] + slotContact[k]] = make_int2( step, k );
}

// copy of a step's snapshot with every contact's friction and restitution replaced by its
// world's values (params: friction, restitution per world)
__global__ void k_patch_params( const DevContact* __restrict__ src, DevContact* __restrict__ dst, const int* __restrict__ slotWorld,
								int slotCount, const float* __restrict__ params )
{
	int k = blockIdx.x * blockDim.x + threadIdx.x;
	if ( k >= slotCount || slotWorld[k] < 0 )
		return;
	DevContact c = src[k];
	c.friction = params[2 * slotWorld[k]];
	c.restitution = params[2 * slotWorld[k] + 1];
	dst[k] = c;
}

// poses between the API's rows of 7 floats (center x, y, rotation c, s, velocity x, y, angular
// velocity) and StepPose records
__global__ void k_poses_from_rows( const float* __restrict__ rows, StepPose* __restrict__ poses, int n )
{
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if ( i >= n )
		return;
	const float* e = rows + 7 * (size_t)i;
	poses[i] = StepPose{ e[0], e[1], e[2], e[3], e[4], e[5], e[6], 0.0f };
}

__global__ void k_rows_from_poses( const StepPose* __restrict__ poses, float* __restrict__ rows, int n )
{
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if ( i >= n )
		return;
	StepPose p = poses[i];
	float* o = rows + 7 * (size_t)i;
	o[0] = p.cx, o[1] = p.cy, o[2] = p.qc, o[3] = p.qs, o[4] = p.vx, o[5] = p.vy, o[6] = p.w;
}

__global__ void k_fill_zero( float* __restrict__ p, size_t n )
{
	size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if ( i < n )
		p[i] = 0.0f;
}

} // namespace b2mi

// ---- host launchers ----
namespace b2mi
{

static inline int blocksFor( size_t n, int block )
{
	return (int)( ( n + block - 1 ) / block );
}

int adjoint_scratch_floats( int bodyCount, int slotCount, int subSteps )
{
	return adjointScratchFloats( bodyCount, slotCount, subSteps );
}

hipError_t launch_snapshot( const DevWorld* worlds, const WorkItem* items, int itemCount, const int* slotContact, DevContact* snap,
							hipStream_t stream )
{
	if ( itemCount <= 0 )
		return hipSuccess;
	hipLaunchKernelGGL( k_snapshot, dim3( itemCount ), dim3( 64 ), 0, stream, worlds, items, slotContact, snap );
	return hipGetLastError();
}

hipError_t launch_gather_poses( const DevWorld* worlds, const int* simBase, const int* simCount, int worldCount, int maxSims,
								StepPose* out, hipStream_t stream )
{
	if ( worldCount <= 0 || maxSims <= 0 )
		return hipSuccess;
	if ( worldCount > 65535 )
		return hipErrorInvalidValue;
	hipLaunchKernelGGL( k_gather_poses, dim3( blocksFor( maxSims, 64 ), worldCount ), dim3( 64 ), 0, stream, worlds, simBase, simCount,
						worldCount, out );
	return hipGetLastError();
}

hipError_t launch_replay_step( const DevWorld* worlds, const WorkItem* items, int itemCount, int blockSize, const int* bodyList,
							   const int2* slotBodies, const DevContact* snap, const StepPose* poseIn, StepPose* poseOut,
							   const int* simBase, float* cf, int cfCapacity, float4* priv, float2* mass, int storeLdsBytes,
							   hipStream_t stream )
{
	if ( itemCount <= 0 )
		return hipSuccess;
	if ( blockSize < 64 || blockSize > kAdjointMaxBlock || storeLdsBytes < 0 || storeLdsBytes > kAdjointStoreLdsBytes )
		return hipErrorInvalidValue;
	if ( storeLdsBytes > 0 )
		hipLaunchKernelGGL( k_replay_step<true>, dim3( itemCount ), dim3( blockSize ), storeLdsBytes, stream, worlds, items, bodyList,
							slotBodies, snap, poseIn, poseOut, simBase, cf, cfCapacity, priv, mass );
	else
		hipLaunchKernelGGL( k_replay_step<false>, dim3( itemCount ), dim3( blockSize ), 0, stream, worlds, items, bodyList, slotBodies,
							snap, poseIn, poseOut, simBase, cf, cfCapacity, priv, mass );
	return hipGetLastError();
}

hipError_t launch_adjoint_step( const DevWorld* worlds, const WorkItem* items, int itemCount, int blockSize, const int* bodyList,
								const int2* slotBodies, const DevContact* snap, const StepPose* poseIn, StepPose* adj,
								const int* simBase, const float4* lamOut, float* lamPrev, const int* prevSlot, float* paramGrad,
								float* cf, int cfCapacity, float4* priv, float2* mass, float* scratch, const int* scratchOffset,
								int storeLdsBytes, hipStream_t stream )
{
	if ( itemCount <= 0 )
		return hipSuccess;
	if ( blockSize < 64 || blockSize > kAdjointMaxBlock || storeLdsBytes < 0 || storeLdsBytes > kAdjointStoreLdsBytes )
		return hipErrorInvalidValue;
	if ( storeLdsBytes > 0 )
		hipLaunchKernelGGL( k_adjoint_step<true>, dim3( itemCount ), dim3( blockSize ), storeLdsBytes, stream, worlds, items, bodyList, slotBodies, snap,
							poseIn, adj, simBase, lamOut, lamPrev, prevSlot, paramGrad, cf, cfCapacity, priv, mass, scratch, scratchOffset );
	else
		hipLaunchKernelGGL( k_adjoint_step<false>, dim3( itemCount ), dim3( blockSize ), 0, stream, worlds, items, bodyList, slotBodies, snap,
							poseIn, adj, simBase, lamOut, lamPrev, prevSlot, paramGrad, cf, cfCapacity, priv, mass, scratch, scratchOffset );
	return hipGetLastError();
}

hipError_t launch_link_slots( const int* slotContact, const int* slotWorld, int slotCount, const int* contactBase,
							  const int2* slotOfContact, int prevStep, int* prevSlot, hipStream_t stream )
{
	if ( slotCount <= 0 )
		return hipSuccess;
	hipLaunchKernelGGL( k_link_slots, dim3( blocksFor( slotCount, 256 ) ), dim3( 256 ), 0, stream, slotContact, slotWorld, slotCount,
						contactBase, slotOfContact, prevStep, prevSlot );
	return hipGetLastError();
}

hipError_t launch_mark_slots( const int* slotContact, const int* slotWorld, int slotCount, const int* contactBase, int step,
							  int2* slotOfContact, hipStream_t stream )
{
	if ( slotCount <= 0 )
		return hipSuccess;
	hipLaunchKernelGGL( k_mark_slots, dim3( blocksFor( slotCount, 256 ) ), dim3( 256 ), 0, stream, slotContact, slotWorld, slotCount,
						contactBase, step, slotOfContact );
	return hipGetLastError();
}

hipError_t launch_patch_params( const DevContact* src, DevContact* dst, const int* slotWorld, int slotCount, const float* params,
								hipStream_t stream )
{
	if ( slotCount <= 0 )
		return hipSuccess;
	hipLaunchKernelGGL( k_patch_params, dim3( blocksFor( slotCount, 256 ) ), dim3( 256 ), 0, stream, src, dst, slotWorld, slotCount,
						params );
	return hipGetLastError();
}

hipError_t launch_poses_from_rows( const float* rows, StepPose* poses, int n, hipStream_t stream )
{
	if ( n <= 0 )
		return hipSuccess;
	hipLaunchKernelGGL( k_poses_from_rows, dim3( blocksFor( n, 256 ) ), dim3( 256 ), 0, stream, rows, poses, n );
	return hipGetLastError();
}

hipError_t launch_rows_from_poses( const StepPose* poses, float* rows, int n, hipStream_t stream )
{
	if ( n <= 0 )
		return hipSuccess;
	hipLaunchKernelGGL( k_rows_from_poses, dim3( blocksFor( n, 256 ) ), dim3( 256 ), 0, stream, poses, rows, n );
	return hipGetLastError();
}

hipError_t launch_fill_zero( float* p, size_t n, hipStream_t stream )
{
	if ( n == 0 )
		return hipSuccess;
	hipLaunchKernelGGL( k_fill_zero, dim3( blocksFor( n, 256 ) ), dim3( 256 ), 0, stream, p, n );
	return hipGetLastError();
}

} // namespace b2mi
