// CPU ORACLE — test infrastructure only (see or_internal.h).
// Step driver (src/world.c:676-776) and the solver of src/solver.c:34-351, 1108-2005 and
// src/contact_solver.c:21-1293. The graph-colour path is the reference's 8-wide AVX2 path
// evaluated one lane at a time: every lane operation is a single IEEE float op, so a
// scalar evaluation in the same order reproduces it bit for bit. Padding lanes of the
// 8-wide layout touch only the identity body and are skipped. The overflow colour uses the
// reference's separate scalar formulas.
#include "or_internal.h"

#include <float.h>
#include <math.h>

typedef struct orPoint
{
	real anchorAx, anchorAy, anchorBx, anchorBy;
	real baseSeparation;
	real relativeVelocity;
	real normalImpulse;
	real tangentImpulse;
	real maxNormalImpulse;
	real normalMass;
	real tangentMass;
} orPoint;

typedef struct orConstraint
{
	int indexA, indexB;
	orContactSim* contact; // where impulses are stored back
	orPoint p[2];
	real nx, ny;
	real invMassA, invMassB, invIA, invIB;
	real friction, restitution;
	real biasRate, massScale, impulseScale;
	int pointCount;
} orConstraint;

// a / b written as maxf(a, b) = a > b ? a : b and minf = a < b ? a : b, the operand order of
// x86 maxps/minps (returns the second operand on ties and NaN), which the reference uses
static inline real or_max( real a, real b )
{
	return a > b ? a : b;
}
static inline real or_min( real a, real b )
{
	return a < b ? a : b;
}

// src/solver.h:144-156
static orSoftness or_make_soft( real hertz, real zeta, real h )
{
	if ( hertz == 0 )
		return ( orSoftness ){ 0, 1, 0 };
	real omega = (real)2 * (real)b2_pi * hertz;
	real a1 = (real)2 * zeta + h * omega;
	real a2 = h * omega * a1;
	real a3 = (real)1 / ( (real)1 + a2 );
	return ( orSoftness ){ omega / a1, a2 * a3, a3 };
}

orSoftness or_make_softness( real hertz, real zeta, real h )
{
	return or_make_soft( hertz, zeta, h );
}

static const orBodyState or_identity_state = { { 0, 0 }, 0, 0, { 0, 0 }, { 1, 0 } };

// ---- prepare (contact_solver.c:597-808 colour lanes; 21-147 overflow) ----
static void or_prepare( orConstraint* c, orContactSim* cs, const orBodyState* states, const orStepContext* ctx,
						real warmStartScale, bool overflow )
{
	const b2Manifold* m = &cs->manifold;
	int indexA = cs->bodySimIndexA, indexB = cs->bodySimIndexB;
	memset( c, 0, sizeof( *c ) );
	c->indexA = indexA;
	c->indexB = indexB;
	c->contact = cs;
	c->pointCount = m->pointCount;

	orVec2r vA = { 0, 0 }, vB = { 0, 0 };
	real wA = 0, wB = 0;
	real mA = cs->invMassA, iA = cs->invIA, mB = cs->invMassB, iB = cs->invIB;
	if ( indexA != OR_NULL )
	{
		vA = states[indexA].v;
		wA = states[indexA].w;
	}
	if ( indexB != OR_NULL )
	{
		vB = states[indexB].v;
		wB = states[indexB].w;
	}
	c->invMassA = mA;
	c->invMassB = mB;
	c->invIA = iA;
	c->invIB = iB;
	orSoftness soft = ( indexA == OR_NULL || indexB == OR_NULL ) ? ctx->staticSoftness : ctx->contactSoftness;
	c->biasRate = soft.biasRate;
	c->massScale = soft.massScale;
	c->impulseScale = soft.impulseScale;
	real nx = m->normal.x, ny = m->normal.y;
	c->nx = nx;
	c->ny = ny;
	c->friction = cs->friction;
	c->restitution = cs->restitution;
	// b2RightPerp(normal)
	real tx = ny, ty = -nx;

	int count = overflow ? m->pointCount : 2;
	for ( int j = 0; j < count; ++j )
	{
		orPoint* cp = c->p + j;
		if ( j >= m->pointCount )
			break; // lane point 2 of a one-point manifold stays zero
		const b2ManifoldPoint* mp = m->points + j;
		real rAx = mp->anchorA.x, rAy = mp->anchorA.y, rBx = mp->anchorB.x, rBy = mp->anchorB.y;
		cp->anchorAx = rAx;
		cp->anchorAy = rAy;
		cp->anchorBx = rBx;
		cp->anchorBy = rBy;
		// separation - dot(rB - rA, normal)
		real dx = rBx - rAx, dy = rBy - rAy;
		cp->baseSeparation = (real)mp->separation - ( dx * nx + dy * ny );
		cp->normalImpulse = warmStartScale * (real)mp->normalImpulse;
		cp->tangentImpulse = warmStartScale * (real)mp->tangentImpulse;
		cp->maxNormalImpulse = 0;
		real rnA = rAx * ny - rAy * nx;
		real rnB = rBx * ny - rBy * nx;
		real kNormal = mA + mB + iA * rnA * rnA + iB * rnB * rnB;
		cp->normalMass = kNormal > 0 ? (real)1 / kNormal : 0;
		real rtA = rAx * ty - rAy * tx;
		real rtB = rBx * ty - rBy * tx;
		real kTangent = mA + mB + iA * rtA * rtA + iB * rtB * rtB;
		cp->tangentMass = kTangent > 0 ? (real)1 / kTangent : 0;
		// vr = v + cross(w, r) = (v.x + (-w) * r.y, v.y + w * r.x)
		real vrAx = vA.x + ( -wA ) * rAy, vrAy = vA.y + wA * rAx;
		real vrBx = vB.x + ( -wB ) * rBy, vrBy = vB.y + wB * rBx;
		cp->relativeVelocity = nx * ( vrBx - vrAx ) + ny * ( vrBy - vrAy );
	}
}

static inline orBodyState* or_state_ptr( orBodyState* states, int index, orBodyState* dummy )
{
	if ( index == OR_NULL )
	{
		*dummy = or_identity_state;
		return dummy;
	}
	return states + index;
}

// ---- graph-colour lane stages (contact_solver.c:810-1178) ----
static void or_lane_apply( orBodyState* bA, orBodyState* bB, const orConstraint* c, real rAx, real rAy, real rBx, real rBy,
						   real Px, real Py )
{
	bA->v.x = bA->v.x - c->invMassA * Px;
	bA->v.y = bA->v.y - c->invMassA * Py;
	bA->w = bA->w - c->invIA * ( rAx * Py - rAy * Px );
	bB->v.x = bB->v.x + c->invMassB * Px;
	bB->v.y = bB->v.y + c->invMassB * Py;
	bB->w = bB->w + c->invIB * ( rBx * Py - rBy * Px );
}

static void or_lane_warm_start( orConstraint* c, orBodyState* states )
{
	orBodyState dA, dB;
	orBodyState bA = *or_state_ptr( states, c->indexA, &dA );
	orBodyState bB = *or_state_ptr( states, c->indexB, &dB );
	real tx = c->ny, ty = (real)0 - c->nx;
	for ( int j = 0; j < 2; ++j )
	{
		const orPoint* cp = c->p + j;
		real Px = cp->normalImpulse * c->nx + cp->tangentImpulse * tx;
		real Py = cp->normalImpulse * c->ny + cp->tangentImpulse * ty;
		bA.w = bA.w - c->invIA * ( cp->anchorAx * Py - cp->anchorAy * Px );
		bA.v.x = bA.v.x - c->invMassA * Px;
		bA.v.y = bA.v.y - c->invMassA * Py;
		bB.w = bB.w + c->invIB * ( cp->anchorBx * Py - cp->anchorBy * Px );
		bB.v.x = bB.v.x + c->invMassB * Px;
		bB.v.y = bB.v.y + c->invMassB * Py;
	}
	if ( c->indexA != OR_NULL )
		states[c->indexA] = bA;
	if ( c->indexB != OR_NULL )
		states[c->indexB] = bB;
}

static void or_lane_solve( orConstraint* c, orBodyState* states, real inv_h, real minBiasVel, bool useBias )
{
	orBodyState dA, dB;
	orBodyState bA = *or_state_ptr( states, c->indexA, &dA );
	orBodyState bB = *or_state_ptr( states, c->indexB, &dB );
	real biasRate = useBias ? c->biasRate : 0;
	real massScale = useBias ? c->massScale : 1;
	real impulseScale = useBias ? c->impulseScale : 0;
	real dpx = bB.dp.x - bA.dp.x, dpy = bB.dp.y - bA.dp.y;

	for ( int j = 0; j < 2; ++j )
	{
		orPoint* cp = c->p + j;
		// moving anchors for the current separation
		real rsAx = bA.dq.x * cp->anchorAx - bA.dq.y * cp->anchorAy;
		real rsAy = bA.dq.y * cp->anchorAx + bA.dq.x * cp->anchorAy;
		real rsBx = bB.dq.x * cp->anchorBx - bB.dq.y * cp->anchorBy;
		real rsBy = bB.dq.y * cp->anchorBx + bB.dq.x * cp->anchorBy;
		real dsx = dpx + ( rsBx - rsAx ), dsy = dpy + ( rsBy - rsAy );
		real s = ( c->nx * dsx + c->ny * dsy ) + cp->baseSeparation;
		real specBias = s * inv_h;
		real softBias = or_max( biasRate * s, minBiasVel );
		real bias = s > 0 ? specBias : softBias;
		real rAx = cp->anchorAx, rAy = cp->anchorAy, rBx = cp->anchorBx, rBy = cp->anchorBy;
		real dvx = ( bB.v.x - bB.w * rBy ) - ( bA.v.x - bA.w * rAy );
		real dvy = ( bB.v.y + bB.w * rBx ) - ( bA.v.y + bA.w * rAx );
		real vn = dvx * c->nx + dvy * c->ny;
		real negImpulse = cp->normalMass * ( massScale * ( vn + bias ) ) + impulseScale * cp->normalImpulse;
		real newImpulse = or_max( cp->normalImpulse - negImpulse, 0 );
		real impulse = newImpulse - cp->normalImpulse;
		cp->normalImpulse = newImpulse;
		cp->maxNormalImpulse = or_max( cp->maxNormalImpulse, newImpulse );
		or_lane_apply( &bA, &bB, c, rAx, rAy, rBx, rBy, impulse * c->nx, impulse * c->ny );
	}

	real tx = c->ny, ty = (real)0 - c->nx;
	for ( int j = 0; j < 2; ++j )
	{
		orPoint* cp = c->p + j;
		real rAx = cp->anchorAx, rAy = cp->anchorAy, rBx = cp->anchorBx, rBy = cp->anchorBy;
		real dvx = ( bB.v.x - bB.w * rBy ) - ( bA.v.x - bA.w * rAy );
		real dvy = ( bB.v.y + bB.w * rBx ) - ( bA.v.y + bA.w * rAx );
		real vt = dvx * tx + dvy * ty;
		real negImpulse = cp->tangentMass * vt;
		real maxFriction = c->friction * cp->normalImpulse;
		real newImpulse = cp->tangentImpulse - negImpulse;
		newImpulse = or_max( (real)0 - maxFriction, or_min( newImpulse, maxFriction ) );
		real impulse = newImpulse - cp->tangentImpulse;
		cp->tangentImpulse = newImpulse;
		or_lane_apply( &bA, &bB, c, rAx, rAy, rBx, rBy, impulse * tx, impulse * ty );
	}

	if ( c->indexA != OR_NULL )
		states[c->indexA] = bA;
	if ( c->indexB != OR_NULL )
		states[c->indexB] = bB;
}

static void or_lane_restitution( orConstraint* c, orBodyState* states, real threshold )
{
	orBodyState dA, dB;
	orBodyState bA = *or_state_ptr( states, c->indexA, &dA );
	orBodyState bB = *or_state_ptr( states, c->indexB, &dB );
	for ( int j = 0; j < 2; ++j )
	{
		orPoint* cp = c->p + j;
		bool skip = ( cp->relativeVelocity + threshold ) > 0 || cp->maxNormalImpulse == 0;
		real mass = skip ? 0 : cp->normalMass;
		real rAx = cp->anchorAx, rAy = cp->anchorAy, rBx = cp->anchorBx, rBy = cp->anchorBy;
		real dvx = ( bB.v.x - bB.w * rBy ) - ( bA.v.x - bA.w * rAy );
		real dvy = ( bB.v.y + bB.w * rBx ) - ( bA.v.y + bA.w * rAx );
		real vn = dvx * c->nx + dvy * c->ny;
		real negImpulse = mass * ( vn + c->restitution * cp->relativeVelocity );
		real newImpulse = or_max( cp->normalImpulse - negImpulse, 0 );
		real impulse = newImpulse - cp->normalImpulse;
		cp->normalImpulse = newImpulse;
		or_lane_apply( &bA, &bB, c, rAx, rAy, rBx, rBy, impulse * c->nx, impulse * c->ny );
	}
	if ( c->indexA != OR_NULL )
		states[c->indexA] = bA;
	if ( c->indexB != OR_NULL )
		states[c->indexB] = bB;
}

// contact_solver.c:1180-1293 (both manifold points, also for one-point manifolds)
static void or_lane_store( const orConstraint* c )
{
	b2Manifold* m = &c->contact->manifold;
	for ( int j = 0; j < 2; ++j )
	{
		m->points[j].normalImpulse = (float)c->p[j].normalImpulse;
		m->points[j].tangentImpulse = (float)c->p[j].tangentImpulse;
		m->points[j].maxNormalImpulse = (float)c->p[j].maxNormalImpulse;
		m->points[j].normalVelocity = (float)c->p[j].relativeVelocity;
	}
}

// ---- overflow stages (contact_solver.c:149-471) ----
static void or_overflow_warm_start( orConstraint* c, orBodyState* states )
{
	orBodyState dA, dB;
	orBodyState* sA = or_state_ptr( states, c->indexA, &dA );
	orBodyState* sB = or_state_ptr( states, c->indexB, &dB );
	orVec2r vA = sA->v, vB = sB->v;
	real wA = sA->w, wB = sB->w;
	real tx = c->ny, ty = -c->nx;
	for ( int j = 0; j < c->pointCount; ++j )
	{
		const orPoint* cp = c->p + j;
		real Px = cp->normalImpulse * c->nx + cp->tangentImpulse * tx;
		real Py = cp->normalImpulse * c->ny + cp->tangentImpulse * ty;
		wA -= c->invIA * ( cp->anchorAx * Py - cp->anchorAy * Px );
		vA.x = vA.x + ( -c->invMassA ) * Px;
		vA.y = vA.y + ( -c->invMassA ) * Py;
		wB += c->invIB * ( cp->anchorBx * Py - cp->anchorBy * Px );
		vB.x = vB.x + c->invMassB * Px;
		vB.y = vB.y + c->invMassB * Py;
	}
	sA->v = vA;
	sA->w = wA;
	sB->v = vB;
	sB->w = wB;
}

static void or_overflow_solve( orConstraint* c, orBodyState* states, real inv_h, real pushout, bool useBias )
{
	orBodyState dA, dB;
	orBodyState* sA = or_state_ptr( states, c->indexA, &dA );
	orBodyState* sB = or_state_ptr( states, c->indexB, &dB );
	orVec2r vA = sA->v, vB = sB->v;
	real wA = sA->w, wB = sB->w;
	orVec2r dqA = sA->dq, dqB = sB->dq;
	real dpx = sB->dp.x - sA->dp.x, dpy = sB->dp.y - sA->dp.y;
	real nx = c->nx, ny = c->ny, tx = ny, ty = -nx;
	real mA = c->invMassA, iA = c->invIA, mB = c->invMassB, iB = c->invIB;

	for ( int j = 0; j < c->pointCount; ++j )
	{
		orPoint* cp = c->p + j;
		real rBqx = dqB.x * cp->anchorBx - dqB.y * cp->anchorBy, rBqy = dqB.y * cp->anchorBx + dqB.x * cp->anchorBy;
		real rAqx = dqA.x * cp->anchorAx - dqA.y * cp->anchorAy, rAqy = dqA.y * cp->anchorAx + dqA.x * cp->anchorAy;
		real dsx = dpx + ( rBqx - rAqx ), dsy = dpy + ( rBqy - rAqy );
		real s = ( dsx * nx + dsy * ny ) + cp->baseSeparation;
		real velocityBias = 0, massScale = 1, impulseScale = 0;
		if ( s > 0 )
			velocityBias = s * inv_h;
		else if ( useBias )
		{
			velocityBias = or_max( c->biasRate * s, -pushout );
			massScale = c->massScale;
			impulseScale = c->impulseScale;
		}
		real rAx = cp->anchorAx, rAy = cp->anchorAy, rBx = cp->anchorBx, rBy = cp->anchorBy;
		real vrAx = vA.x + ( -wA ) * rAy, vrAy = vA.y + wA * rAx;
		real vrBx = vB.x + ( -wB ) * rBy, vrBy = vB.y + wB * rBx;
		real vn = ( vrBx - vrAx ) * nx + ( vrBy - vrAy ) * ny;
		real impulse = -cp->normalMass * massScale * ( vn + velocityBias ) - impulseScale * cp->normalImpulse;
		real newImpulse = or_max( cp->normalImpulse + impulse, 0 );
		impulse = newImpulse - cp->normalImpulse;
		cp->normalImpulse = newImpulse;
		cp->maxNormalImpulse = or_max( cp->maxNormalImpulse, impulse );
		real Px = impulse * nx, Py = impulse * ny;
		vA.x = vA.x - mA * Px;
		vA.y = vA.y - mA * Py;
		wA -= iA * ( rAx * Py - rAy * Px );
		vB.x = vB.x + mB * Px;
		vB.y = vB.y + mB * Py;
		wB += iB * ( rBx * Py - rBy * Px );
	}

	for ( int j = 0; j < c->pointCount; ++j )
	{
		orPoint* cp = c->p + j;
		real rAx = cp->anchorAx, rAy = cp->anchorAy, rBx = cp->anchorBx, rBy = cp->anchorBy;
		real vrBx = vB.x + ( -wB ) * rBy, vrBy = vB.y + wB * rBx;
		real vrAx = vA.x + ( -wA ) * rAy, vrAy = vA.y + wA * rAx;
		real vt = ( vrBx - vrAx ) * tx + ( vrBy - vrAy ) * ty;
		real impulse = cp->tangentMass * ( -vt );
		real maxFriction = c->friction * cp->normalImpulse;
		real sum = cp->tangentImpulse + impulse;
		real newImpulse = sum < -maxFriction ? -maxFriction : ( sum > maxFriction ? maxFriction : sum );
		impulse = newImpulse - cp->tangentImpulse;
		cp->tangentImpulse = newImpulse;
		real Px = impulse * tx, Py = impulse * ty;
		vA.x = vA.x - mA * Px;
		vA.y = vA.y - mA * Py;
		wA -= iA * ( rAx * Py - rAy * Px );
		vB.x = vB.x + mB * Px;
		vB.y = vB.y + mB * Py;
		wB += iB * ( rBx * Py - rBy * Px );
	}
	sA->v = vA;
	sA->w = wA;
	sB->v = vB;
	sB->w = wB;
}

static void or_overflow_restitution( orConstraint* c, orBodyState* states, real threshold )
{
	if ( c->restitution == 0 )
		return;
	orBodyState dA, dB;
	orBodyState* sA = or_state_ptr( states, c->indexA, &dA );
	orBodyState* sB = or_state_ptr( states, c->indexB, &dB );
	orVec2r vA = sA->v, vB = sB->v;
	real wA = sA->w, wB = sB->w;
	real nx = c->nx, ny = c->ny;
	real mA = c->invMassA, iA = c->invIA, mB = c->invMassB, iB = c->invIB;
	for ( int j = 0; j < c->pointCount; ++j )
	{
		orPoint* cp = c->p + j;
		if ( cp->relativeVelocity > -threshold || cp->maxNormalImpulse == 0 )
			continue;
		real rAx = cp->anchorAx, rAy = cp->anchorAy, rBx = cp->anchorBx, rBy = cp->anchorBy;
		real vrBx = vB.x + ( -wB ) * rBy, vrBy = vB.y + wB * rBx;
		real vrAx = vA.x + ( -wA ) * rAy, vrAy = vA.y + wA * rAx;
		real vn = ( vrBx - vrAx ) * nx + ( vrBy - vrAy ) * ny;
		real impulse = -cp->normalMass * ( vn + c->restitution * cp->relativeVelocity );
		real newImpulse = or_max( cp->normalImpulse + impulse, 0 );
		impulse = newImpulse - cp->normalImpulse;
		cp->normalImpulse = newImpulse;
		cp->maxNormalImpulse = or_max( cp->maxNormalImpulse, impulse );
		real Px = impulse * nx, Py = impulse * ny;
		vA.x = vA.x - mA * Px;
		vA.y = vA.y - mA * Py;
		wA -= iA * ( rAx * Py - rAy * Px );
		vB.x = vB.x + mB * Px;
		vB.y = vB.y + mB * Py;
		wB += iB * ( rBx * Py - rBy * Px );
	}
	sA->v = vA;
	sA->w = wA;
	sB->v = vB;
	sB->w = wB;
}

static void or_overflow_store( const orConstraint* c )
{
	b2Manifold* m = &c->contact->manifold;
	for ( int j = 0; j < m->pointCount; ++j )
	{
		m->points[j].normalImpulse = (float)c->p[j].normalImpulse;
		m->points[j].tangentImpulse = (float)c->p[j].tangentImpulse;
		m->points[j].maxNormalImpulse = (float)c->p[j].maxNormalImpulse;
		m->points[j].normalVelocity = (float)c->p[j].relativeVelocity;
	}
}

// ---- body stages (solver.c:34-163) ----
static void or_integrate_velocities( orWorld* world, const orStepContext* ctx )
{
	real h = ctx->h;
	real gx = world->gravity.x, gy = world->gravity.y;
	real maxLinearSpeed = (real)OR_MAX_TRANSLATION * (real)ctx->inv_dt;
	real maxAngularSpeed = (real)OR_MAX_ROTATION * (real)ctx->inv_dt;
	real maxLinearSpeedSquared = maxLinearSpeed * maxLinearSpeed;
	real maxAngularSpeedSquared = maxAngularSpeed * maxAngularSpeed;
	for ( int i = 0; i < world->awakeSims.count; ++i )
	{
		orBodySim* sim = world->awakeSims.data + i;
		orBodyState* state = world->awakeStates.data + i;
		orVec2r v = state->v;
		real w = state->w;
		real linearDamping = (real)1 / ( (real)1 + h * (real)sim->linearDamping );
		real angularDamping = (real)1 / ( (real)1 + h * (real)sim->angularDamping );
		// h * invMass * (force + mass * gravityScale * gravity)
		real mg = (real)sim->mass * (real)sim->gravityScale;
		real fx = (real)sim->force.x + mg * gx, fy = (real)sim->force.y + mg * gy;
		real hm = h * (real)sim->invMass;
		real lvdx = hm * fx, lvdy = hm * fy;
		real avd = h * (real)sim->invInertia * (real)sim->torque;
		v.x = lvdx + linearDamping * v.x;
		v.y = lvdy + linearDamping * v.y;
		w = avd + angularDamping * w;
		if ( v.x * v.x + v.y * v.y > maxLinearSpeedSquared )
		{
#if ORACLE_SOLVER_F64
			real len = sqrt( v.x * v.x + v.y * v.y );
#else
			real len = sqrtf( v.x * v.x + v.y * v.y );
#endif
			real ratio = maxLinearSpeed / len;
			v.x = ratio * v.x;
			v.y = ratio * v.y;
			sim->isSpeedCapped = true;
		}
		if ( w * w > maxAngularSpeedSquared )
		{
			real ratio = maxAngularSpeed / ( w < 0 ? -w : w );
			w *= ratio;
			sim->isSpeedCapped = true;
		}
		state->v = v;
		state->w = w;
	}
}

static void or_integrate_positions( orWorld* world, const orStepContext* ctx )
{
	real h = ctx->h;
	for ( int i = 0; i < world->awakeStates.count; ++i )
	{
		orBodyState* s = world->awakeStates.data + i;
		// b2IntegrateRotation(dq, h * w)
		real da = h * s->w;
		real c2 = s->dq.x - da * s->dq.y;
		real s2 = s->dq.y + da * s->dq.x;
#if ORACLE_SOLVER_F64
		real mag = sqrt( s2 * s2 + c2 * c2 );
#else
		real mag = sqrtf( s2 * s2 + c2 * c2 );
#endif
		real inv = mag > 0 ? (real)1 / mag : 0;
		s->dq.x = c2 * inv;
		s->dq.y = s2 * inv;
		s->dp.x = s->dp.x + h * s->v.x;
		s->dp.y = s->dp.y + h * s->v.y;
	}
}

static b2AABB or_expand( b2AABB a, float d )
{
	a.lowerBound.x -= d;
	a.lowerBound.y -= d;
	a.upperBound.x += d;
	a.upperBound.y += d;
	return a;
}

// solver.c:165-351
static void or_finalize_bodies( orWorld* world, const orStepContext* ctx, orBitSet* enlargedSims, orIntArray* fastBodies,
								orIntArray* bulletBodies )
{
	bool enableSleep = world->enableSleep;
	bool enableContinuous = world->enableContinuous;
	real timeStep = ctx->dt;
	real invTimeStep = ctx->inv_dt;
	for ( int simIndex = 0; simIndex < world->awakeSims.count; ++simIndex )
	{
		orBodyState* state = world->awakeStates.data + simIndex;
		orBodySim* sim = world->awakeSims.data + simIndex;
		orBody* body = world->bodies.data + sim->bodyId;
		orVec2r v = state->v;
		real w = state->w;
#if ORACLE_SOLVER_F64
		sim->centerR.x = sim->centerR.x + state->dp.x;
		sim->centerR.y = sim->centerR.y + state->dp.y;
		{
			// q = normalize(dq * q)
			real qs = state->dq.y * sim->qR.x + state->dq.x * sim->qR.y;
			real qc = state->dq.x * sim->qR.x - state->dq.y * sim->qR.y;
			real mag = sqrt( qs * qs + qc * qc );
			real inv = mag > 0 ? 1.0 / mag : 0.0;
			sim->qR = ( orVec2r ){ qc * inv, qs * inv };
		}
		sim->center = ( b2Vec2 ){ (float)sim->centerR.x, (float)sim->centerR.y };
		sim->transform.q = ( b2Rot ){ (float)sim->qR.x, (float)sim->qR.y };
		real maxVelocity = sqrt( v.x * v.x + v.y * v.y ) + fabs( w ) * sim->maxExtent;
		real maxDeltaPosition = sqrt( state->dp.x * state->dp.x + state->dp.y * state->dp.y ) + fabs( state->dq.y ) * sim->maxExtent;
#else
		sim->center = b2Add( sim->center, ( b2Vec2 ){ state->dp.x, state->dp.y } );
		sim->transform.q = b2NormalizeRot( b2MulRot( ( b2Rot ){ state->dq.x, state->dq.y }, sim->transform.q ) );
		real maxVelocity = b2Length( ( b2Vec2 ){ v.x, v.y } ) + b2AbsFloat( w ) * sim->maxExtent;
		real maxDeltaPosition = b2Length( ( b2Vec2 ){ state->dp.x, state->dp.y } ) + b2AbsFloat( state->dq.y ) * sim->maxExtent;
#endif
		real positionSleepFactor = 0.5f;
		real sleepVelocity = or_max( maxVelocity, positionSleepFactor * invTimeStep * maxDeltaPosition );
		state->dp = ( orVec2r ){ 0, 0 };
		state->dq = ( orVec2r ){ 1, 0 };
#if ORACLE_SOLVER_F64
		{
			real lx = sim->localCenter.x, ly = sim->localCenter.y;
			real px = sim->centerR.x - ( sim->qR.x * lx - sim->qR.y * ly );
			real py = sim->centerR.y - ( sim->qR.y * lx + sim->qR.x * ly );
			sim->transform.p = ( b2Vec2 ){ (float)px, (float)py };
		}
#else
		sim->transform.p = b2Sub( sim->center, b2RotateVector( sim->transform.q, sim->localCenter ) );
#endif
		// solver.c:226-231
		b2BodyMoveEvent* moveEvent = world->moveEvents.data + simIndex;
		moveEvent->transform = sim->transform;
		moveEvent->bodyId = ( b2BodyId ){ sim->bodyId + 1, (uint16_t)world->worldId, body->revision };
		moveEvent->userData = body->userData;
		moveEvent->fellAsleep = false;

		sim->force = b2Vec2_zero;
		sim->torque = 0.0f;
		body->isSpeedCapped = sim->isSpeedCapped;
		sim->isSpeedCapped = false;
		sim->isFast = false;

		if ( enableSleep == false || body->enableSleep == false || sleepVelocity > body->sleepThreshold )
		{
			const real safetyFactor = 0.5f;
			if ( body->type == b2_dynamicBody && enableContinuous && maxVelocity * timeStep > safetyFactor * sim->minExtent )
			{
				if ( sim->isBullet )
					OR_PUSH( *bulletBodies, simIndex );
				else
					OR_PUSH( *fastBodies, simIndex );
				sim->isFast = true;
			}
			else
			{
				sim->center0 = sim->center;
				sim->rotation0 = sim->transform.q;
			}
		}
		else
		{
			sim->center0 = sim->center;
			sim->rotation0 = sim->transform.q;
		}

		b2Transform transform = sim->transform;
		for ( int id = body->headShapeId; id != OR_NULL; id = world->shapes.data[id].nextShapeId )
		{
			orShape* shape = world->shapes.data + id;
			if ( sim->isFast )
			{
				shape->isFast = true;
				or_bitset_set( enlargedSims, simIndex );
			}
			else
			{
				b2AABB aabb = or_expand( or_compute_shape_aabb( shape, transform ), OR_SPECULATIVE );
				shape->aabb = aabb;
				if ( b2AABB_Contains( shape->fatAABB, aabb ) == false )
				{
					shape->fatAABB = or_expand( aabb, OR_AABB_MARGIN );
					shape->enlargedAABB = true;
					or_bitset_set( enlargedSims, simIndex );
				}
			}
		}
	}
}

static void or_buffer_move( orWorld* world, int proxyKey )
{
	if ( or_bitset_get( &world->moveSet, (uint32_t)proxyKey ) == false )
	{
		or_bitset_set( &world->moveSet, (uint32_t)proxyKey );
		OR_PUSH( world->moveArray, proxyKey );
	}
}

// src/solver.c:786-794
typedef struct orContinuousQuery
{
	orWorld* world;
	orBodySim* fastSim;
	orShape* fastShape;
	b2Vec2 centroid1, centroid2;
	b2Sweep sweep;
	float fraction;
} orContinuousQuery;

// b2MakeSweep (src/body.h:159-168)
static b2Sweep or_make_sweep( const orBodySim* sim )
{
	b2Sweep s;
	s.c1 = sim->center0;
	s.c2 = sim->center;
	s.q1 = sim->rotation0;
	s.q2 = sim->transform.q;
	s.localCenter = sim->localCenter;
	return s;
}

// b2MakeShapeDistanceProxy (src/shape.c:713-734)
static b2DistanceProxy or_shape_proxy( const orShape* shape )
{
	switch ( shape->type )
	{
		case b2_capsuleShape:
			return b2MakeProxy( &shape->capsule.center1, 2, shape->capsule.radius );
		case b2_circleShape:
			return b2MakeProxy( &shape->circle.center, 1, shape->circle.radius );
		case b2_polygonShape:
			return b2MakeProxy( shape->polygon.vertices, shape->polygon.count, shape->polygon.radius );
		case b2_segmentShape:
			return b2MakeProxy( &shape->segment.point1, 2, 0.0f );
		default:
		{
			b2DistanceProxy empty = { 0 };
			return empty;
		}
	}
}

// b2GetShapeCentroid (src/shape.c:480-497)
static b2Vec2 or_fast_centroid( const orShape* s )
{
	switch ( s->type )
	{
		case b2_capsuleShape:
			return b2Lerp( s->capsule.center1, s->capsule.center2, 0.5f );
		case b2_circleShape:
			return s->circle.center;
		case b2_polygonShape:
			return s->polygon.centroid;
		case b2_segmentShape:
			return b2Lerp( s->segment.point1, s->segment.point2, 0.5f );
		default:
			return b2Vec2_zero;
	}
}

// src/solver.c:796-901: the time of impact against one candidate shape, keeping the smallest
// fraction (sensors, custom filters and smooth segments are not part of this restatement)
static bool or_continuous_callback( int proxyId, int shapeId, void* context )
{
	(void)proxyId;
	orContinuousQuery* q = context;
	orShape* fastShape = q->fastShape;
	if ( shapeId == fastShape->id )
		return true;
	orWorld* world = q->world;
	orShape* shape = world->shapes.data + shapeId;
	if ( shape->bodyId == fastShape->bodyId )
		return true;
	b2Filter a = fastShape->filter, b = shape->filter;
	bool collide = ( a.groupIndex == b.groupIndex && a.groupIndex != 0 ) ? a.groupIndex > 0
																		  : ( ( a.maskBits & b.categoryBits ) != 0 && ( a.categoryBits & b.maskBits ) != 0 );
	if ( collide == false )
		return true;
	orBody* body = world->bodies.data + shape->bodyId;
	orBodySim* bodySim = or_get_body_sim( world, body );
	OR_ASSERT( body->type == b2_staticBody || q->fastSim->isBullet );
	if ( bodySim->isBullet )
		return true;
	orBody* fastBody = world->bodies.data + q->fastSim->bodyId;
	if ( or_should_bodies_collide( world, fastBody, body ) == false )
		return true;

	b2TOIInput input;
	input.proxyA = or_shape_proxy( shape );
	input.proxyB = or_shape_proxy( fastShape );
	input.sweepA = or_make_sweep( bodySim );
	input.sweepB = q->sweep;
	input.tMax = q->fraction;
	b2TOIOutput output = b2TimeOfImpact( &input );
	if ( 0.0f < output.t && output.t < q->fraction )
		q->fraction = output.t;
	else if ( 0.0f == output.t )
	{
		// a small circle at the fast shape's centroid
		b2Vec2 centroid = or_fast_centroid( fastShape );
		input.proxyB = b2MakeProxy( &centroid, 1, OR_SPECULATIVE );
		output = b2TimeOfImpact( &input );
		if ( 0.0f < output.t && output.t < q->fraction )
			q->fraction = output.t;
	}
	return true;
}

// src/solver.c:904-1066: the fast body's sweep against the static tree (and, for a bullet, the
// kinematic and dynamic trees); on a hit the body moves back to the time of impact
static void or_solve_continuous( orWorld* world, int simIndex )
{
	orBodySim* sim = world->awakeSims.data + simIndex;
	orBody* body = world->bodies.data + sim->bodyId;
	b2Sweep sweep = or_make_sweep( sim );
	b2Transform xf1 = { b2Sub( sweep.c1, b2RotateVector( sweep.q1, sweep.localCenter ) ), sweep.q1 };
	b2Transform xf2 = { b2Sub( sweep.c2, b2RotateVector( sweep.q2, sweep.localCenter ) ), sweep.q2 };
	orContinuousQuery q;
	q.world = world;
	q.fastSim = sim;
	q.sweep = sweep;
	q.fraction = 1.0f;
	for ( int id = body->headShapeId; id != OR_NULL; )
	{
		orShape* fastShape = world->shapes.data + id;
		id = fastShape->nextShapeId;
		fastShape->isFast = false;
		q.fastShape = fastShape;
		q.centroid1 = b2TransformPoint( xf1, fastShape->localCentroid );
		q.centroid2 = b2TransformPoint( xf2, fastShape->localCentroid );
		b2AABB box1 = fastShape->aabb;
		b2AABB box2 = or_compute_shape_aabb( fastShape, xf2 );
		b2AABB box = b2AABB_Union( box1, box2 );
		fastShape->aabb = box2;
		if ( fastShape->isSensor )
			continue;
		or_tree_query( world->trees + b2_staticBody, box, b2_defaultMaskBits, or_continuous_callback, &q );
		if ( sim->isBullet )
		{
			or_tree_query( world->trees + b2_kinematicBody, box, b2_defaultMaskBits, or_continuous_callback, &q );
			or_tree_query( world->trees + b2_dynamicBody, box, b2_defaultMaskBits, or_continuous_callback, &q );
		}
	}

	if ( q.fraction < 1.0f )
	{
		b2Rot rq = b2NLerp( sweep.q1, sweep.q2, q.fraction );
		b2Vec2 c = b2Lerp( sweep.c1, sweep.c2, q.fraction );
		b2Vec2 origin = b2Sub( c, b2RotateVector( rq, sweep.localCenter ) );
		b2Transform transform = { origin, rq };
		sim->transform = transform;
		sim->center = c;
		sim->rotation0 = rq;
		sim->center0 = c;
#if ORACLE_SOLVER_F64
		sim->centerR = ( orVec2r ){ c.x, c.y };
		sim->qR = ( orVec2r ){ rq.c, rq.s };
#endif
		world->toiHits += 1;
		for ( int id = body->headShapeId; id != OR_NULL; id = world->shapes.data[id].nextShapeId )
		{
			orShape* shape = world->shapes.data + id;
			b2AABB aabb = or_expand( or_compute_shape_aabb( shape, transform ), OR_SPECULATIVE );
			shape->aabb = aabb;
			if ( b2AABB_Contains( shape->fatAABB, aabb ) == false )
			{
				shape->fatAABB = or_expand( aabb, OR_AABB_MARGIN );
				shape->enlargedAABB = true;
				sim->enlargeAABB = true;
			}
		}
	}
	else
	{
		sim->rotation0 = sim->transform.q;
		sim->center0 = sim->center;
		for ( int id = body->headShapeId; id != OR_NULL; id = world->shapes.data[id].nextShapeId )
		{
			orShape* shape = world->shapes.data + id;
			if ( b2AABB_Contains( shape->fatAABB, shape->aabb ) == false )
			{
				shape->fatAABB = or_expand( shape->aabb, OR_AABB_MARGIN );
				shape->enlargedAABB = true;
				sim->enlargeAABB = true;
			}
		}
	}
}

// src/solver.c:1826-1873 (and 1889-1937 for bullets): the fast shapes whose fat AABB grew,
// enlarged in the dynamic tree (they are already in the move buffer)
static void or_enlarge_fast( orWorld* world, const orIntArray* list )
{
	for ( int i = 0; i < list->count; ++i )
	{
		orBodySim* sim = world->awakeSims.data + list->data[i];
		if ( sim->enlargeAABB == false )
			continue;
		sim->enlargeAABB = false;
		orBody* body = world->bodies.data + sim->bodyId;
		for ( int id = body->headShapeId; id != OR_NULL; id = world->shapes.data[id].nextShapeId )
		{
			orShape* shape = world->shapes.data + id;
			if ( shape->enlargedAABB == false )
				continue;
			shape->enlargedAABB = false;
			or_tree_enlarge_proxy( world->trees + b2_dynamicBody, shape->proxyKey >> 2, shape->fatAABB );
		}
	}
}

void or_solve( orWorld* world, orStepContext* ctx )
{
	world->stepIndex += 1;
	int awakeBodyCount = world->awakeSims.count;
	if ( awakeBodyCount == 0 )
		return;

	orBodyState* states = world->awakeStates.data;
	real warmStartScale = world->enableWarmStarting ? 1 : 0;

	// active colours: those with contacts or joints (solver.c:1145-1171)
	int activeColors[OR_GRAPH_COLORS];
	int activeColorCount = 0, total = 0;
	for ( int i = 0; i < OR_OVERFLOW; ++i )
	{
		if ( world->colors[i].contacts.count + world->colors[i].joints.count > 0 )
		{
			activeColors[activeColorCount++] = i;
			total += world->colors[i].contacts.count;
		}
	}

	// prepare joints: the colours' (b2_stagePrepareJoints) then the overflow's
	// (solver.c:598-618); each writes only its own joint
	for ( int c = 0; c < OR_GRAPH_COLORS; ++c )
	{
		orJointSimArray* joints = &world->colors[c].joints;
		for ( int i = 0; i < joints->count; ++i )
			or_prepare_joint( world, joints->data + i, ctx );
	}
	orGraphColor* overflow = world->colors + OR_OVERFLOW;
	int overflowCount = overflow->contacts.count;
	orConstraint* cons = malloc( ( total + overflowCount + 1 ) * sizeof( orConstraint ) );
	int colorStart[OR_GRAPH_COLORS + 1];
	int k = 0;
	for ( int a = 0; a < activeColorCount; ++a )
	{
		orGraphColor* color = world->colors + activeColors[a];
		colorStart[a] = k;
		for ( int i = 0; i < color->contacts.count; ++i )
			or_prepare( cons + k++, color->contacts.data + i, states, ctx, warmStartScale, false );
	}
	colorStart[activeColorCount] = k;
	orConstraint* ovf = cons + k;
	for ( int i = 0; i < overflowCount; ++i )
		or_prepare( ovf + i, overflow->contacts.data + i, states, ctx, warmStartScale, true );

	real inv_h = ctx->inv_h;
	real pushout = world->contactPushoutVelocity;
	real minBiasVel = -pushout;
	orJointSimArray* overflowJoints = &overflow->joints;
	for ( int sub = 0; sub < ctx->subStepCount; ++sub )
	{
		or_integrate_velocities( world, ctx );
		// warm start: overflow joints and contacts always (solver.c:638-640), the colours' joint
		// and contact blocks when warm starting is on (solver.c:398-410)
		for ( int i = 0; i < overflowJoints->count; ++i )
			or_warm_start_joint( overflowJoints->data + i, states );
		for ( int i = 0; i < overflowCount; ++i )
			or_overflow_warm_start( ovf + i, states );
		if ( world->enableWarmStarting )
			for ( int a = 0; a < activeColorCount; ++a )
			{
				orJointSimArray* joints = &world->colors[activeColors[a]].joints;
				for ( int i = 0; i < joints->count; ++i )
					or_warm_start_joint( joints->data + i, states );
				for ( int i = colorStart[a]; i < colorStart[a + 1]; ++i )
					or_lane_warm_start( cons + i, states );
			}
		for ( int i = 0; i < overflowJoints->count; ++i )
			or_solve_joint( overflowJoints->data + i, states, ctx, true );
		for ( int i = 0; i < overflowCount; ++i )
			or_overflow_solve( ovf + i, states, inv_h, pushout, true );
		for ( int a = 0; a < activeColorCount; ++a )
		{
			orJointSimArray* joints = &world->colors[activeColors[a]].joints;
			for ( int i = 0; i < joints->count; ++i )
				or_solve_joint( joints->data + i, states, ctx, true );
			for ( int i = colorStart[a]; i < colorStart[a + 1]; ++i )
				or_lane_solve( cons + i, states, inv_h, minBiasVel, true );
		}
		or_integrate_positions( world, ctx );
		for ( int i = 0; i < overflowJoints->count; ++i )
			or_solve_joint( overflowJoints->data + i, states, ctx, false );
		for ( int i = 0; i < overflowCount; ++i )
			or_overflow_solve( ovf + i, states, inv_h, pushout, false );
		for ( int a = 0; a < activeColorCount; ++a )
		{
			orJointSimArray* joints = &world->colors[activeColors[a]].joints;
			for ( int i = 0; i < joints->count; ++i )
				or_solve_joint( joints->data + i, states, ctx, false );
			for ( int i = colorStart[a]; i < colorStart[a + 1]; ++i )
				or_lane_solve( cons + i, states, inv_h, minBiasVel, false );
		}
	}
	real threshold = ctx->restitutionThreshold;
	for ( int i = 0; i < overflowCount; ++i )
		or_overflow_restitution( ovf + i, states, threshold );
	for ( int i = 0; i < colorStart[activeColorCount]; ++i )
		or_lane_restitution( cons + i, states, threshold );
	for ( int i = 0; i < overflowCount; ++i )
		or_overflow_store( ovf + i );
	for ( int i = 0; i < colorStart[activeColorCount]; ++i )
		or_lane_store( cons + i );
	free( cons );

	// finalize (solver.c:1628-1652), one move event per awake body (solver.c:1170-1172)
	OR_RESERVE( world->moveEvents, awakeBodyCount );
	world->moveEvents.count = awakeBodyCount;
	orBitSet enlargedSims = { 0 };
	or_bitset_set_count_and_clear( &enlargedSims, (uint32_t)awakeBodyCount );
	orIntArray fastBodies = { 0 }, bulletBodies = { 0 };
	or_finalize_bodies( world, ctx, &enlargedSims, &fastBodies, &bulletBodies );

	// hit events (solver.c:1670-1723): every colour, overflow included, in local order; the
	// manifold point of the fastest approach above the threshold that received an impulse
	for ( int c = 0; c < OR_GRAPH_COLORS; ++c )
	{
		const orGraphColor* color = world->colors + c;
		for ( int i = 0; i < color->contacts.count; ++i )
		{
			const orContactSim* cs = color->contacts.data + i;
			if ( ( cs->simFlags & OR_SIM_HIT_EVENT ) == 0 )
				continue;
			b2ContactHitEvent event = { 0 };
			event.approachSpeed = world->hitEventThreshold;
			bool hit = false;
			for ( int p = 0; p < cs->manifold.pointCount; ++p )
			{
				const b2ManifoldPoint* mp = cs->manifold.points + p;
				float approachSpeed = -mp->normalVelocity;
				if ( approachSpeed > event.approachSpeed && mp->normalImpulse > 0.0f )
				{
					event.approachSpeed = approachSpeed;
					event.point = mp->point;
					hit = true;
				}
			}
			if ( hit )
			{
				event.normal = cs->manifold.normal;
				const orShape* shapeA = world->shapes.data + cs->shapeIdA;
				const orShape* shapeB = world->shapes.data + cs->shapeIdB;
				event.shapeIdA = ( b2ShapeId ){ shapeA->id + 1, (uint16_t)world->worldId, shapeA->revision };
				event.shapeIdB = ( b2ShapeId ){ shapeB->id + 1, (uint16_t)world->worldId, shapeB->revision };
				OR_PUSH( world->contactHit, event );
			}
		}
	}

	// enlarge proxies in sim order, building the next move array (solver.c:1741-1797)
	for ( uint32_t w = 0; w < enlargedSims.blockCount; ++w )
	{
		uint64_t word = enlargedSims.bits[w];
		while ( word != 0 )
		{
			int simIndex = (int)( 64 * w + __builtin_ctzll( word ) );
			orBody* body = world->bodies.data + world->awakeSims.data[simIndex].bodyId;
			for ( int id = body->headShapeId; id != OR_NULL; id = world->shapes.data[id].nextShapeId )
			{
				orShape* shape = world->shapes.data + id;
				if ( shape->enlargedAABB )
				{
					int type = shape->proxyKey & 3;
					or_tree_enlarge_proxy( world->trees + type, shape->proxyKey >> 2, shape->fatAABB );
					or_buffer_move( world, shape->proxyKey );
					shape->enlargedAABB = false;
				}
				else if ( shape->isFast )
				{
					or_buffer_move( world, shape->proxyKey );
				}
			}
			word &= word - 1;
		}
	}

	// continuous (solver.c:1807-1937): the fast bodies, their shapes enlarged in the dynamic tree,
	// then the bullets (which see the fast bodies' final poses), theirs enlarged
	for ( int i = 0; i < fastBodies.count; ++i )
		or_solve_continuous( world, fastBodies.data[i] );
	or_enlarge_fast( world, &fastBodies );
	for ( int i = 0; i < bulletBodies.count; ++i )
		or_solve_continuous( world, bulletBodies.data[i] );
	or_enlarge_fast( world, &bulletBodies );
	OR_FREE( fastBodies );
	OR_FREE( bulletBodies );
	or_bitset_free( &enlargedSims );
}

// src/world.c:676-776
B2_API void b2World_Step( b2WorldId worldId, float timeStep, int subStepCount )
{
	orWorld* world = or_get_world( worldId );
	if ( world->locked )
		return;
	// the last step's events go (world.c:685-692)
	world->moveEvents.count = 0;
	world->sensorBegin.count = 0;
	world->sensorEnd.count = 0;
	world->contactBegin.count = 0;
	world->contactEnd.count = 0;
	world->contactHit.count = 0;
	memset( &world->profile, 0, sizeof( world->profile ) );
	if ( timeStep == 0.0f )
		return;
	world->locked = true;

	or_update_pairs( world );

	orStepContext ctx = { 0 };
	ctx.dt = timeStep;
	ctx.subStepCount = b2MaxInt( 1, subStepCount );
	if ( timeStep > 0.0f )
	{
		ctx.inv_dt = 1.0f / timeStep;
		ctx.h = timeStep / ctx.subStepCount;
		ctx.inv_h = ctx.subStepCount * ctx.inv_dt;
	}
	float contactHertz = b2MinFloat( world->contactHertz, 0.25f * ctx.inv_h );
#if ORACLE_SOLVER_F64
	// the double build keeps the float step sizes and derives softness in double
	ctx.contactSoftness = or_make_soft( contactHertz, world->contactDampingRatio, ctx.h );
	ctx.staticSoftness = or_make_soft( 2.0 * (double)contactHertz, world->contactDampingRatio, ctx.h );
	ctx.jointSoftness = or_make_soft( b2MinFloat( world->jointHertz, 0.125f * ctx.inv_h ), world->jointDampingRatio, ctx.h );
#else
	ctx.contactSoftness = or_make_soft( contactHertz, world->contactDampingRatio, ctx.h );
	ctx.staticSoftness = or_make_soft( 2.0f * contactHertz, world->contactDampingRatio, ctx.h );
	ctx.jointSoftness = or_make_soft( b2MinFloat( world->jointHertz, 0.125f * ctx.inv_h ), world->jointDampingRatio, ctx.h );
#endif
	world->inv_h = ctx.inv_h;
	ctx.restitutionThreshold = world->restitutionThreshold;
	ctx.enableWarmStarting = world->enableWarmStarting;

	or_collide( world, &ctx );
	if ( ctx.dt > 0.0f )
		or_solve( world, &ctx );
	world->locked = false;
}
