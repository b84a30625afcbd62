// CPU ORACLE — test infrastructure only (see or_internal.h).
// Joints: definitions and creation (src/joint.c:19-694), destruction (696-812), the generic and
// revolute queries and setters, the pair filter they add (src/body.c:1724-1760), and the solver
// functions of each joint type (src/revolute_joint.c, src/weld_joint.c), called by or_solve in the
// reference's stage order. Solver quantities are `real` (double in the FP64 build, where the
// joint's anchors still come from the float body sims).
#include "or_internal.h"

#include <float.h>
#include <math.h>

#define OR_JOINT( jointId ) \
	orWorld* world = or_get_world_index( ( jointId ).world0 ); \
	OR_ASSERT( world != NULL ); \
	orJoint* joint = or_get_joint_full( world, jointId ); \
	(void)joint

// ---- definitions (src/joint.c:19-80) ----

B2_API b2DistanceJointDef b2DefaultDistanceJointDef( void )
{
	b2DistanceJointDef def = { 0 };
	def.length = 1.0f;
	def.maxLength = OR_HUGE;
	def.internalValue = OR_COOKIE;
	return def;
}

B2_API b2MotorJointDef b2DefaultMotorJointDef( void )
{
	b2MotorJointDef def = { 0 };
	def.maxForce = 1.0f;
	def.maxTorque = 1.0f;
	def.correctionFactor = 0.3f;
	def.internalValue = OR_COOKIE;
	return def;
}

B2_API b2MouseJointDef b2DefaultMouseJointDef( void )
{
	b2MouseJointDef def = { 0 };
	def.hertz = 4.0f;
	def.dampingRatio = 1.0f;
	def.maxForce = 1.0f;
	def.internalValue = OR_COOKIE;
	return def;
}

B2_API b2PrismaticJointDef b2DefaultPrismaticJointDef( void )
{
	b2PrismaticJointDef def = { 0 };
	def.localAxisA = ( b2Vec2 ){ 1.0f, 0.0f };
	def.internalValue = OR_COOKIE;
	return def;
}

B2_API b2RevoluteJointDef b2DefaultRevoluteJointDef( void )
{
	b2RevoluteJointDef def = { 0 };
	def.drawSize = 0.25f;
	def.internalValue = OR_COOKIE;
	return def;
}

B2_API b2WeldJointDef b2DefaultWeldJointDef( void )
{
	b2WeldJointDef def = { 0 };
	def.internalValue = OR_COOKIE;
	return def;
}

B2_API b2WheelJointDef b2DefaultWheelJointDef( void )
{
	b2WheelJointDef def = { 0 };
	def.localAxisA.y = 1.0f;
	def.enableSpring = true;
	def.hertz = 1.0f;
	def.dampingRatio = 0.7f;
	def.internalValue = OR_COOKIE;
	return def;
}

// ---- lookup (src/joint.c:82-128) ----

static orJoint* or_get_joint_full( orWorld* world, b2JointId jointId )
{
	int id = jointId.index1 - 1;
	OR_ASSERT( 0 <= id && id < world->joints.count );
	orJoint* joint = world->joints.data + id;
	OR_ASSERT( joint->jointId == id && joint->revision == jointId.revision );
	return joint;
}

static orJointSim* or_get_joint_sim( orWorld* world, orJoint* joint )
{
	if ( joint->setIndex == 2 )
		return world->colors[joint->colorIndex].joints.data + joint->localIndex;
	OR_ASSERT( joint->setIndex == 0 );
	return world->staticJoints.data + joint->localIndex;
}

static orJointSim* or_joint_sim_of_type( b2JointId jointId, b2JointType type )
{
	OR_JOINT( jointId );
	OR_ASSERT( world->locked == false );
	OR_ASSERT( joint->type == type );
	return or_get_joint_sim( world, joint );
}

// ---- colouring (src/constraint_graph.c:216-305) ----

static int or_assign_joint_color( orWorld* world, int bodyIdA, int bodyIdB, bool staticA, bool staticB )
{
	OR_ASSERT( staticA == false || staticB == false );
	for ( int i = 0; i < OR_OVERFLOW; ++i )
	{
		orBitSet* set = &world->colors[i].bodySet;
		if ( staticA == false && staticB == false )
		{
			if ( or_bitset_get( set, bodyIdA ) || or_bitset_get( set, bodyIdB ) )
				continue;
			or_bitset_set( set, bodyIdA );
			or_bitset_set( set, bodyIdB );
			return i;
		}
		if ( staticA == false )
		{
			if ( or_bitset_get( set, bodyIdA ) )
				continue;
			or_bitset_set( set, bodyIdA );
			return i;
		}
		if ( or_bitset_get( set, bodyIdB ) )
			continue;
		or_bitset_set( set, bodyIdB );
		return i;
	}
	return OR_OVERFLOW;
}

// ---- creation (src/joint.c:130-281) ----

static orJointSim* or_create_joint( orWorld* world, orBody* bodyA, orBody* bodyB, void* userData, float drawSize, b2JointType type,
									bool collideConnected, b2JointId* outId )
{
	OR_ASSERT( world->locked == false );
	// the reference links a joint of two non-dynamic awake bodies into no island (src/island.c:315)
	OR_ASSERT( ( bodyA->setIndex == 0 && bodyB->setIndex == 0 ) || bodyA->type == b2_dynamicBody ||
			   bodyB->type == b2_dynamicBody );
	int jointId = or_alloc_id( &world->jointIdPool );
	if ( jointId == world->joints.count )
	{
		orJoint empty = { 0 };
		OR_PUSH( world->joints, empty );
	}
	orJoint* joint = world->joints.data + jointId;
	joint->jointId = jointId;
	joint->userData = userData;
	joint->revision += 1;
	joint->setIndex = OR_NULL;
	joint->colorIndex = OR_NULL;
	joint->localIndex = OR_NULL;
	joint->drawSize = drawSize;
	joint->type = type;
	joint->collideConnected = collideConnected;

	// doubly linked lists on body A then body B, head insertion
	for ( int e = 0; e < 2; ++e )
	{
		orBody* body = e == 0 ? bodyA : bodyB;
		joint->edges[e].bodyId = body->id;
		joint->edges[e].prevKey = OR_NULL;
		joint->edges[e].nextKey = body->headJointKey;
		int key = ( jointId << 1 ) | e;
		if ( body->headJointKey != OR_NULL )
		{
			orJoint* other = world->joints.data + ( body->headJointKey >> 1 );
			other->edges[body->headJointKey & 1].prevKey = key;
		}
		body->headJointKey = key;
		body->jointCount += 1;
	}

	orJointSim* sim;
	if ( bodyA->setIndex == 0 && bodyB->setIndex == 0 )
	{
		joint->setIndex = 0;
		joint->localIndex = world->staticJoints.count;
		orJointSim empty = { 0 };
		OR_PUSH( world->staticJoints, empty );
		sim = world->staticJoints.data + joint->localIndex;
	}
	else
	{
		joint->setIndex = 2;
		int colorIndex = or_assign_joint_color( world, bodyA->id, bodyB->id, bodyA->setIndex == 0, bodyB->setIndex == 0 );
		orGraphColor* color = world->colors + colorIndex;
		joint->colorIndex = colorIndex;
		joint->localIndex = color->joints.count;
		orJointSim empty = { 0 };
		OR_PUSH( color->joints, empty );
		sim = color->joints.data + joint->localIndex;
	}
	sim->jointId = jointId;
	sim->bodyIdA = bodyA->id;
	sim->bodyIdB = bodyB->id;
	sim->type = type;
	*outId = ( b2JointId ){ jointId + 1, (uint16_t)world->worldId, joint->revision };
	return sim;
}

// src/joint.c:283-318
static void or_destroy_contacts_between_bodies( orWorld* world, orBody* bodyA, orBody* bodyB )
{
	int contactKey, otherBodyId;
	if ( bodyA->contactCount < bodyB->contactCount )
	{
		contactKey = bodyA->headContactKey;
		otherBodyId = bodyB->id;
	}
	else
	{
		contactKey = bodyB->headContactKey;
		otherBodyId = bodyA->id;
	}
	while ( contactKey != OR_NULL )
	{
		orContact* contact = world->contacts.data + ( contactKey >> 1 );
		int edgeIndex = contactKey & 1;
		contactKey = contact->edges[edgeIndex].nextKey;
		if ( contact->edges[edgeIndex ^ 1].bodyId == otherBodyId )
			or_destroy_contact( world, contact );
	}
}

#define OR_JOINT_BODIES( def ) \
	orWorld* world = or_get_world( worldId ); \
	OR_ASSERT( ( def )->internalValue == OR_COOKIE ); \
	orBody* bodyA = or_get_body_full( world, ( def )->bodyIdA ); \
	orBody* bodyB = or_get_body_full( world, ( def )->bodyIdB ); \
	b2JointId jointId

#define OR_FINISH_JOINT( def ) \
	do \
	{ \
		if ( ( def )->collideConnected == false ) \
			or_destroy_contacts_between_bodies( world, bodyA, bodyB ); \
		return jointId; \
	} while ( 0 )

// src/joint.c:320-368
B2_API b2JointId b2CreateDistanceJoint( b2WorldId worldId, const b2DistanceJointDef* def )
{
	OR_JOINT_BODIES( def );
	OR_ASSERT( def->length > 0.0f );
	orJointSim* j = or_create_joint( world, bodyA, bodyB, def->userData, 1.0f, b2_distanceJoint, def->collideConnected, &jointId );
	j->localOriginAnchorA = def->localAnchorA;
	j->localOriginAnchorB = def->localAnchorB;
	j->length = b2MaxFloat( def->length, OR_LINEAR_SLOP );
	j->hertz = def->hertz;
	j->dampingRatio = def->dampingRatio;
	j->lower = b2MaxFloat( def->minLength, OR_LINEAR_SLOP );
	j->upper = b2MaxFloat( def->minLength, def->maxLength );
	j->maxMotor = def->maxMotorForce;
	j->motorSpeed = def->motorSpeed;
	j->enableSpring = def->enableSpring;
	j->enableLimit = def->enableLimit;
	j->enableMotor = def->enableMotor;
	OR_FINISH_JOINT( def );
}

// src/joint.c:370-405
B2_API b2JointId b2CreateMotorJoint( b2WorldId worldId, const b2MotorJointDef* def )
{
	OR_JOINT_BODIES( def );
	orJointSim* j = or_create_joint( world, bodyA, bodyB, def->userData, 1.0f, b2_motorJoint, def->collideConnected, &jointId );
	j->localOriginAnchorA = ( b2Vec2 ){ 0.0f, 0.0f };
	j->localOriginAnchorB = ( b2Vec2 ){ 0.0f, 0.0f };
	j->linearOffset = def->linearOffset;
	j->angularOffset = def->angularOffset;
	j->maxForce = def->maxForce;
	j->maxTorque = def->maxTorque;
	j->correctionFactor = b2ClampFloat( def->correctionFactor, 0.0f, 1.0f );
	OR_FINISH_JOINT( def );
}

// src/joint.c:407-439 (the mouse joint leaves contacts between its bodies alone)
B2_API b2JointId b2CreateMouseJoint( b2WorldId worldId, const b2MouseJointDef* def )
{
	OR_JOINT_BODIES( def );
	b2Transform transformA = b2Body_GetTransform( def->bodyIdA );
	b2Transform transformB = b2Body_GetTransform( def->bodyIdB );
	orJointSim* j = or_create_joint( world, bodyA, bodyB, def->userData, 1.0f, b2_mouseJoint, def->collideConnected, &jointId );
	j->localOriginAnchorA = b2InvTransformPoint( transformA, def->target );
	j->localOriginAnchorB = b2InvTransformPoint( transformB, def->target );
	j->target = def->target;
	j->hertz = def->hertz;
	j->dampingRatio = def->dampingRatio;
	j->maxForce = def->maxForce;
	return jointId;
}

// src/joint.c:441-492
B2_API b2JointId b2CreateRevoluteJoint( b2WorldId worldId, const b2RevoluteJointDef* def )
{
	OR_JOINT_BODIES( def );
	orJointSim* j =
		or_create_joint( world, bodyA, bodyB, def->userData, def->drawSize, b2_revoluteJoint, def->collideConnected, &jointId );
	j->localOriginAnchorA = def->localAnchorA;
	j->localOriginAnchorB = def->localAnchorB;
	j->referenceAngle = b2ClampFloat( def->referenceAngle, -b2_pi, b2_pi );
	j->hertz = def->hertz;
	j->dampingRatio = def->dampingRatio;
	j->lower = b2MinFloat( def->lowerAngle, def->upperAngle );
	j->upper = b2MaxFloat( def->lowerAngle, def->upperAngle );
	j->lower = b2ClampFloat( j->lower, -b2_pi, b2_pi );
	j->upper = b2ClampFloat( j->upper, -b2_pi, b2_pi );
	j->maxMotor = def->maxMotorTorque;
	j->motorSpeed = def->motorSpeed;
	j->enableSpring = def->enableSpring;
	j->enableLimit = def->enableLimit;
	j->enableMotor = def->enableMotor;
	OR_FINISH_JOINT( def );
}

// src/joint.c:494-563
B2_API b2JointId b2CreatePrismaticJoint( b2WorldId worldId, const b2PrismaticJointDef* def )
{
	OR_JOINT_BODIES( def );
	orJointSim* j = or_create_joint( world, bodyA, bodyB, def->userData, 1.0f, b2_prismaticJoint, def->collideConnected, &jointId );
	j->localOriginAnchorA = def->localAnchorA;
	j->localOriginAnchorB = def->localAnchorB;
	j->localAxisA = b2Normalize( def->localAxisA );
	j->referenceAngle = def->referenceAngle;
	j->hertz = def->hertz;
	j->dampingRatio = def->dampingRatio;
	j->lower = def->lowerTranslation;
	j->upper = def->upperTranslation;
	j->maxMotor = def->maxMotorForce;
	j->motorSpeed = def->motorSpeed;
	j->enableSpring = def->enableSpring;
	j->enableLimit = def->enableLimit;
	j->enableMotor = def->enableMotor;
	OR_FINISH_JOINT( def );
}

// src/joint.c:565-603
B2_API b2JointId b2CreateWeldJoint( b2WorldId worldId, const b2WeldJointDef* def )
{
	OR_JOINT_BODIES( def );
	orJointSim* j = or_create_joint( world, bodyA, bodyB, def->userData, 1.0f, b2_weldJoint, def->collideConnected, &jointId );
	j->localOriginAnchorA = def->localAnchorA;
	j->localOriginAnchorB = def->localAnchorB;
	j->referenceAngle = def->referenceAngle;
	j->linearHertz = def->linearHertz;
	j->linearDampingRatio = def->linearDampingRatio;
	j->angularHertz = def->angularHertz;
	j->angularDampingRatio = def->angularDampingRatio;
	OR_FINISH_JOINT( def );
}

// src/joint.c:605-658
B2_API b2JointId b2CreateWheelJoint( b2WorldId worldId, const b2WheelJointDef* def )
{
	OR_JOINT_BODIES( def );
	orJointSim* j = or_create_joint( world, bodyA, bodyB, def->userData, 1.0f, b2_wheelJoint, def->collideConnected, &jointId );
	j->localOriginAnchorA = def->localAnchorA;
	j->localOriginAnchorB = def->localAnchorB;
	j->localAxisA = b2Normalize( def->localAxisA );
	j->lower = def->lowerTranslation;
	j->upper = def->upperTranslation;
	j->maxMotor = def->maxMotorTorque;
	j->motorSpeed = def->motorSpeed;
	j->hertz = def->hertz;
	j->dampingRatio = def->dampingRatio;
	j->enableSpring = def->enableSpring;
	j->enableLimit = def->enableLimit;
	j->enableMotor = def->enableMotor;
	OR_FINISH_JOINT( def );
}

// ---- destruction (src/joint.c:660-756, src/constraint_graph.c:294-324) ----

void or_destroy_joint_internal( orWorld* world, orJoint* joint )
{
	int jointId = joint->jointId;
	for ( int e = 0; e < 2; ++e )
	{
		orJointEdge* edge = joint->edges + e;
		orBody* body = world->bodies.data + edge->bodyId;
		if ( edge->prevKey != OR_NULL )
			world->joints.data[edge->prevKey >> 1].edges[edge->prevKey & 1].nextKey = edge->nextKey;
		if ( edge->nextKey != OR_NULL )
			world->joints.data[edge->nextKey >> 1].edges[edge->nextKey & 1].prevKey = edge->prevKey;
		int key = ( jointId << 1 ) | e;
		if ( body->headJointKey == key )
			body->headJointKey = edge->nextKey;
		body->jointCount -= 1;
	}
	orJointSimArray* array;
	if ( joint->setIndex == 2 )
	{
		orGraphColor* color = world->colors + joint->colorIndex;
		if ( joint->colorIndex != OR_OVERFLOW )
		{
			or_bitset_clear( &color->bodySet, joint->edges[0].bodyId );
			or_bitset_clear( &color->bodySet, joint->edges[1].bodyId );
		}
		array = &color->joints;
	}
	else
		array = &world->staticJoints;
	int last = array->count - 1;
	if ( joint->localIndex < last )
	{
		array->data[joint->localIndex] = array->data[last];
		world->joints.data[array->data[joint->localIndex].jointId].localIndex = joint->localIndex;
	}
	array->count -= 1;
	joint->setIndex = OR_NULL;
	joint->localIndex = OR_NULL;
	joint->colorIndex = OR_NULL;
	joint->jointId = OR_NULL;
	or_free_id( &world->jointIdPool, jointId );
}

B2_API void b2DestroyJoint( b2JointId jointId )
{
	OR_JOINT( jointId );
	OR_ASSERT( world->locked == false );
	or_destroy_joint_internal( world, joint );
}

// ---- queries (src/joint.c:758-981, src/body.c:1700-1722) ----

B2_API bool b2Joint_IsValid( b2JointId id )
{
	orWorld* world = or_get_world_index( id.world0 );
	if ( world == NULL )
		return false;
	int jointId = id.index1 - 1;
	if ( jointId < 0 || world->joints.count <= jointId )
		return false;
	orJoint* joint = world->joints.data + jointId;
	if ( joint->jointId == OR_NULL )
		return false;
	return id.revision == joint->revision;
}

B2_API b2JointType b2Joint_GetType( b2JointId jointId )
{
	OR_JOINT( jointId );
	return joint->type;
}

static b2BodyId or_body_id( orWorld* world, int bodyId )
{
	orBody* body = world->bodies.data + bodyId;
	return ( b2BodyId ){ bodyId + 1, (uint16_t)world->worldId, body->revision };
}

B2_API b2BodyId b2Joint_GetBodyA( b2JointId jointId )
{
	OR_JOINT( jointId );
	return or_body_id( world, joint->edges[0].bodyId );
}

B2_API b2BodyId b2Joint_GetBodyB( b2JointId jointId )
{
	OR_JOINT( jointId );
	return or_body_id( world, joint->edges[1].bodyId );
}

B2_API b2Vec2 b2Joint_GetLocalAnchorA( b2JointId jointId )
{
	OR_JOINT( jointId );
	return or_get_joint_sim( world, joint )->localOriginAnchorA;
}

B2_API b2Vec2 b2Joint_GetLocalAnchorB( b2JointId jointId )
{
	OR_JOINT( jointId );
	return or_get_joint_sim( world, joint )->localOriginAnchorB;
}

B2_API bool b2Joint_GetCollideConnected( b2JointId jointId )
{
	OR_JOINT( jointId );
	return joint->collideConnected;
}

B2_API void b2Joint_SetUserData( b2JointId jointId, void* userData )
{
	OR_JOINT( jointId );
	joint->userData = userData;
}

B2_API void* b2Joint_GetUserData( b2JointId jointId )
{
	OR_JOINT( jointId );
	return joint->userData;
}

B2_API void b2Joint_WakeBodies( b2JointId jointId )
{
	OR_JOINT( jointId );
}

// src/joint.c:897-934 and the per-type force functions
B2_API b2Vec2 b2Joint_GetConstraintForce( b2JointId jointId )
{
	OR_JOINT( jointId );
	orJointSim* j = or_get_joint_sim( world, joint );
	float inv_h = world->inv_h;
	switch ( joint->type )
	{
		case b2_revoluteJoint:
		case b2_weldJoint:
		case b2_motorJoint:
		case b2_mouseJoint:
			return ( b2Vec2 ){ inv_h * (float)j->linearImpulse.x, inv_h * (float)j->linearImpulse.y };
		case b2_distanceJoint:
		{
			// src/distance_joint.c:183-196: along the current anchor axis
			b2Transform xfA = b2Body_GetTransform( or_body_id( world, j->bodyIdA ) );
			b2Transform xfB = b2Body_GetTransform( or_body_id( world, j->bodyIdB ) );
			b2Vec2 pA = b2TransformPoint( xfA, j->localOriginAnchorA );
			b2Vec2 pB = b2TransformPoint( xfB, j->localOriginAnchorB );
			b2Vec2 axis = b2Normalize( b2Sub( pB, pA ) );
			float force = (float)( j->impulse + j->lowerImpulse - j->upperImpulse + j->motorImpulse ) * inv_h;
			return b2MulSV( force, axis );
		}
		case b2_prismaticJoint:
		{
			// src/prismatic_joint.c:144-160
			b2Transform xfA = b2Body_GetTransform( or_body_id( world, j->bodyIdA ) );
			b2Vec2 axisA = b2RotateVector( xfA.q, j->localAxisA );
			b2Vec2 perpA = b2LeftPerp( axisA );
			float perpForce = inv_h * (float)j->linearImpulse.x;
			float axialForce = inv_h * (float)( j->motorImpulse + j->lowerImpulse - j->upperImpulse );
			return b2Add( b2MulSV( perpForce, perpA ), b2MulSV( axialForce, axisA ) );
		}
		case b2_wheelJoint:
		{
			// src/wheel_joint.c:145-158: the axis of the last prepare
			b2Vec2 axisA = { (float)j->axisA.x, (float)j->axisA.y };
			b2Vec2 perpA = b2LeftPerp( axisA );
			float perpForce = inv_h * (float)j->impulse;
			float axialForce = inv_h * (float)( j->springImpulse + j->lowerImpulse - j->upperImpulse );
			return b2Add( b2MulSV( perpForce, perpA ), b2MulSV( axialForce, axisA ) );
		}
		default:
			OR_ASSERT( false );
			return ( b2Vec2 ){ 0.0f, 0.0f };
	}
}

// src/joint.c:943-978 and the per-type torque functions
B2_API float b2Joint_GetConstraintTorque( b2JointId jointId )
{
	OR_JOINT( jointId );
	orJointSim* j = or_get_joint_sim( world, joint );
	float inv_h = world->inv_h;
	switch ( joint->type )
	{
		case b2_revoluteJoint:
			return inv_h * (float)( j->motorImpulse + j->lowerImpulse - j->upperImpulse );
		case b2_weldJoint:
		case b2_motorJoint:
		case b2_mouseJoint:
			return inv_h * (float)j->angularImpulse;
		case b2_distanceJoint:
			return 0.0f;
		case b2_prismaticJoint:
			return inv_h * (float)j->linearImpulse.y;
		case b2_wheelJoint:
			return inv_h * (float)j->motorImpulse;
		default:
			OR_ASSERT( false );
			return 0.0f;
	}
}

B2_API int b2Body_GetJointCount( b2BodyId bodyId )
{
	orWorld* world = or_get_world_index( bodyId.world0 );
	orBody* body = or_get_body_full( world, bodyId );
	return body->jointCount;
}

B2_API int b2Body_GetJoints( b2BodyId bodyId, b2JointId* jointArray, int capacity )
{
	orWorld* world = or_get_world_index( bodyId.world0 );
	orBody* body = or_get_body_full( world, bodyId );
	int jointKey = body->headJointKey, index = 0;
	while ( jointKey != OR_NULL && index < capacity )
	{
		orJoint* joint = world->joints.data + ( jointKey >> 1 );
		jointArray[index++] = ( b2JointId ){ joint->jointId + 1, (uint16_t)world->worldId, joint->revision };
		jointKey = joint->edges[jointKey & 1].nextKey;
	}
	return index;
}

// ---- revolute queries and setters (src/revolute_joint.c:18-168) ----

B2_API void b2RevoluteJoint_EnableMotor( b2JointId jointId, bool enableMotor )
{
	orJointSim* j = or_joint_sim_of_type( jointId, b2_revoluteJoint );
	if ( enableMotor != j->enableMotor )
	{
		j->enableMotor = enableMotor;
		j->motorImpulse = 0;
	}
}

B2_API void b2RevoluteJoint_SetMotorSpeed( b2JointId jointId, float motorSpeed )
{
	or_joint_sim_of_type( jointId, b2_revoluteJoint )->motorSpeed = motorSpeed;
}

B2_API float b2RevoluteJoint_GetMotorTorque( b2JointId jointId )
{
	OR_JOINT( jointId );
	orJointSim* j = or_joint_sim_of_type( jointId, b2_revoluteJoint );
	return world->inv_h * (float)j->motorImpulse;
}

B2_API void b2RevoluteJoint_EnableLimit( b2JointId jointId, bool enableLimit )
{
	orJointSim* j = or_joint_sim_of_type( jointId, b2_revoluteJoint );
	if ( enableLimit != j->enableLimit )
	{
		j->enableLimit = enableLimit;
		j->lowerImpulse = 0;
		j->upperImpulse = 0;
	}
}

B2_API void b2RevoluteJoint_SetLimits( b2JointId jointId, float lower, float upper )
{
	orJointSim* j = or_joint_sim_of_type( jointId, b2_revoluteJoint );
	if ( lower != j->lower || upper != j->upper )
	{
		j->lower = b2MinFloat( lower, upper );
		j->upper = b2MaxFloat( lower, upper );
		j->lowerImpulse = 0;
		j->upperImpulse = 0;
	}
}

B2_API float b2RevoluteJoint_GetAngle( b2JointId jointId )
{
	OR_JOINT( jointId );
	orJointSim* j = or_joint_sim_of_type( jointId, b2_revoluteJoint );
	b2Transform transformA = b2Body_GetTransform( or_body_id( world, j->bodyIdA ) );
	b2Transform transformB = b2Body_GetTransform( or_body_id( world, j->bodyIdB ) );
	// b2RelativeAngle then b2UnwindAngle (include/box2d/math_functions.h:387-409)
	float s = transformB.q.s * transformA.q.c - transformB.q.c * transformA.q.s;
	float c = transformB.q.c * transformA.q.c + transformB.q.s * transformA.q.s;
	float angle = atan2f( s, c ) - j->referenceAngle;
	if ( angle < -b2_pi )
		angle = angle + 2.0f * b2_pi;
	else if ( angle > b2_pi )
		angle = angle - 2.0f * b2_pi;
	return angle;
}

// ---- pair filter (src/body.c:1724-1760) ----

bool or_should_bodies_collide( orWorld* world, orBody* bodyA, orBody* bodyB )
{
	if ( bodyA->type != b2_dynamicBody && bodyB->type != b2_dynamicBody )
		return false;
	int jointKey, otherBodyId;
	if ( bodyA->jointCount < bodyB->jointCount )
	{
		jointKey = bodyA->headJointKey;
		otherBodyId = bodyB->id;
	}
	else
	{
		jointKey = bodyB->headJointKey;
		otherBodyId = bodyA->id;
	}
	while ( jointKey != OR_NULL )
	{
		orJoint* joint = world->joints.data + ( jointKey >> 1 );
		int edgeIndex = jointKey & 1;
		if ( joint->collideConnected == false && joint->edges[edgeIndex ^ 1].bodyId == otherBodyId )
			return false;
		jointKey = joint->edges[edgeIndex].nextKey;
	}
	return true;
}

// ---- solver ----

static const orBodyState or_joint_dummy = { { 0, 0 }, 0, 0, { 0, 0 }, { 1, 0 } };

static inline real or_atan2( real y, real x )
{
#if ORACLE_SOLVER_F64
	return atan2( y, x );
#else
	return atan2f( y, x );
#endif
}

// b2RelativeAngle (include/box2d/math_functions.h:387-394) of rotations (bc, bs) and (ac, as)
static inline real or_relative_angle( real bc, real bs, real ac, real as )
{
	real s = bs * ac - bc * as;
	real c = bc * ac + bs * as;
	return or_atan2( s, c );
}

// b2UnwindAngle (include/box2d/math_functions.h:397-409)
static inline real or_unwind( real angle )
{
	if ( angle < (real)-b2_pi )
		return angle + (real)2.0f * (real)b2_pi;
	if ( angle > (real)b2_pi )
		return angle - (real)2.0f * (real)b2_pi;
	return angle;
}

static inline real or_maxr( real a, real b )
{
	return a > b ? a : b;
}

static inline real or_clampr( real a, real lo, real hi )
{
	return a < lo ? lo : ( a > hi ? hi : a );
}

// b2RotateVector of a float vector by a rotation
static inline orVec2r or_rotate( real c, real s, real vx, real vy )
{
	return ( orVec2r ){ c * vx - s * vy, s * vx + c * vy };
}

static inline orBodyState* or_joint_state( orBodyState* states, int index, orBodyState* dummy )
{
	if ( index == OR_NULL )
	{
		*dummy = or_joint_dummy;
		return dummy;
	}
	return states + index;
}

// b2Solve22 (include/box2d/math_functions.h:489-499)
static inline orVec2r or_solve22( real a11, real a12, real a21, real a22, real bx, real by )
{
	real det = a11 * a22 - a12 * a21;
	if ( det != 0 )
		det = (real)1 / det;
	return ( orVec2r ){ det * ( a22 * bx - a12 * by ), det * ( a11 * by - a21 * bx ) };
}

static inline real or_cross( orVec2r a, orVec2r b )
{
	return a.x * b.y - a.y * b.x;
}

static inline real or_dot( orVec2r a, orVec2r b )
{
	return a.x * b.x + a.y * b.y;
}

// b2Normalize (src/math_functions.c:55-66)
static inline orVec2r or_normalize( orVec2r v )
{
#if ORACLE_SOLVER_F64
	real length = sqrt( v.x * v.x + v.y * v.y );
	if ( length < DBL_EPSILON )
		return ( orVec2r ){ 0, 0 };
#else
	real length = sqrtf( v.x * v.x + v.y * v.y );
	if ( length < FLT_EPSILON )
		return ( orVec2r ){ 0, 0 };
#endif
	real inv = (real)1 / length;
	return ( orVec2r ){ inv * v.x, inv * v.y };
}

static inline real or_length( orVec2r v )
{
#if ORACLE_SOLVER_F64
	return sqrt( v.x * v.x + v.y * v.y );
#else
	return sqrtf( v.x * v.x + v.y * v.y );
#endif
}

static inline orVec2r or_add( orVec2r a, orVec2r b )
{
	return ( orVec2r ){ a.x + b.x, a.y + b.y };
}

static inline orVec2r or_sub( orVec2r a, orVec2r b )
{
	return ( orVec2r ){ a.x - b.x, a.y - b.y };
}

static inline orVec2r or_scale( real s, orVec2r v )
{
	return ( orVec2r ){ s * v.x, s * v.y };
}

// cross(w, r) = (-w r.y, w r.x)
static inline orVec2r or_cross_sv( real w, orVec2r r )
{
	return ( orVec2r ){ -w * r.y, w * r.x };
}

static inline orVec2r or_left_perp( orVec2r v )
{
	return ( orVec2r ){ -v.y, v.x };
}

// b2GetInverse22 (include/box2d/math_functions.h:471-485)
static inline orMat22r or_inverse22( real a, real b, real c, real d )
{
	real det = a * d - b * c;
	if ( det != 0 )
		det = (real)1 / det;
	return ( orMat22r ){ { det * d, -det * c }, { -det * b, det * a } };
}

// b2MulMV (include/box2d/math_functions.h:461-468)
static inline orVec2r or_mul_mv( orMat22r A, orVec2r v )
{
	return ( orVec2r ){ A.cx.x * v.x + A.cy.x * v.y, A.cx.y * v.x + A.cy.y * v.y };
}

// vA -= mA P, wA -= iA LA, vB += mB P, wB += iB LB
static inline void or_apply( orVec2r* vA, real* wA, orVec2r* vB, real* wB, const orJointSim* j, orVec2r P, real LA, real LB )
{
	vA->x = vA->x - j->invMassA * P.x;
	vA->y = vA->y - j->invMassA * P.y;
	*wA -= j->invIA * LA;
	vB->x = vB->x + j->invMassB * P.x;
	vB->y = vB->y + j->invMassB * P.y;
	*wB += j->invIB * LB;
}

static inline orVec2r or_dp( const orBodyState* s )
{
	return s->dp;
}

// the common part of b2Prepare*Joint: body masses and indices (src/revolute_joint.c:201-249)
void or_prepare_joint( orWorld* world, orJointSim* j, const orStepContext* ctx )
{
	orBody* bodyA = world->bodies.data + j->bodyIdA;
	orBody* bodyB = world->bodies.data + j->bodyIdB;
	OR_ASSERT( bodyA->setIndex == 2 || bodyB->setIndex == 2 );
	orBodySim* simA = bodyA->setIndex == 0 ? world->staticSims.data + bodyA->localIndex : world->awakeSims.data + bodyA->localIndex;
	orBodySim* simB = bodyB->setIndex == 0 ? world->staticSims.data + bodyB->localIndex : world->awakeSims.data + bodyB->localIndex;
	real mA = simA->invMass, iA = simA->invInertia, mB = simB->invMass, iB = simB->invInertia;
	j->invMassA = mA;
	j->invMassB = mB;
	j->invIA = iA;
	j->invIB = iB;
	j->indexA = bodyA->setIndex == 2 ? bodyA->localIndex : OR_NULL;
	j->indexB = bodyB->setIndex == 2 ? bodyB->localIndex : OR_NULL;
	b2Rot qA = simA->transform.q, qB = simB->transform.q;
	b2Vec2 lcA = simA->localCenter, lcB = simB->localCenter;
	real h = ctx->h;
	switch ( j->type )
	{
		case b2_revoluteJoint:
		{
			// src/revolute_joint.c:251-268
			j->anchorA = or_rotate( qA.c, qA.s, j->localOriginAnchorA.x - lcA.x, j->localOriginAnchorA.y - lcA.y );
			j->anchorB = or_rotate( qB.c, qB.s, j->localOriginAnchorB.x - lcB.x, j->localOriginAnchorB.y - lcB.y );
			j->deltaCenter = ( orVec2r ){ (real)simB->center.x - (real)simA->center.x, (real)simB->center.y - (real)simA->center.y };
			j->deltaAngle = or_relative_angle( qB.c, qB.s, qA.c, qA.s ) - (real)j->referenceAngle;
			j->deltaAngle = or_unwind( j->deltaAngle );
			real k = iA + iB;
			j->axialMass = k > 0 ? (real)1 / k : 0;
			j->springSoftness = or_make_softness( j->hertz, j->dampingRatio, h );
			if ( ctx->enableWarmStarting == false )
			{
				j->linearImpulse = ( orVec2r ){ 0, 0 };
				j->springImpulse = 0;
				j->motorImpulse = 0;
				j->lowerImpulse = 0;
				j->upperImpulse = 0;
			}
			break;
		}
		case b2_weldJoint:
		{
			// src/weld_joint.c:137-171
			j->anchorA = or_rotate( qA.c, qA.s, j->localOriginAnchorA.x - lcA.x, j->localOriginAnchorA.y - lcA.y );
			j->anchorB = or_rotate( qB.c, qB.s, j->localOriginAnchorB.x - lcB.x, j->localOriginAnchorB.y - lcB.y );
			j->deltaCenter = ( orVec2r ){ (real)simB->center.x - (real)simA->center.x, (real)simB->center.y - (real)simA->center.y };
			j->deltaAngle = or_relative_angle( qB.c, qB.s, qA.c, qA.s ) - (real)j->referenceAngle;
			real ka = iA + iB;
			j->axialMass = ka > 0 ? (real)1 / ka : 0;
			j->linearSoftness = j->linearHertz == 0.0f ? ctx->jointSoftness : or_make_softness( j->linearHertz, j->linearDampingRatio, h );
			j->angularSoftness =
				j->angularHertz == 0.0f ? ctx->jointSoftness : or_make_softness( j->angularHertz, j->angularDampingRatio, h );
			if ( ctx->enableWarmStarting == false )
			{
				j->linearImpulse = ( orVec2r ){ 0, 0 };
				j->angularImpulse = 0;
			}
			break;
		}
		case b2_distanceJoint:
		{
			// src/distance_joint.c:235-258
			j->anchorA = or_rotate( qA.c, qA.s, j->localOriginAnchorA.x - lcA.x, j->localOriginAnchorA.y - lcA.y );
			j->anchorB = or_rotate( qB.c, qB.s, j->localOriginAnchorB.x - lcB.x, j->localOriginAnchorB.y - lcB.y );
			j->deltaCenter = ( orVec2r ){ (real)simB->center.x - (real)simA->center.x, (real)simB->center.y - (real)simA->center.y };
			orVec2r rA = j->anchorA, rB = j->anchorB;
			orVec2r separation = or_add( or_sub( rB, rA ), j->deltaCenter );
			orVec2r axis = or_normalize( separation );
			real crA = or_cross( rA, axis ), crB = or_cross( rB, axis );
			real k = mA + mB + iA * crA * crA + iB * crB * crB;
			j->axialMass = k > 0 ? (real)1 / k : 0;
			j->springSoftness = or_make_softness( j->hertz, j->dampingRatio, h );
			if ( ctx->enableWarmStarting == false )
			{
				j->impulse = 0;
				j->lowerImpulse = 0;
				j->upperImpulse = 0;
				j->motorImpulse = 0;
			}
			break;
		}
		case b2_prismaticJoint:
		{
			// src/prismatic_joint.c:245-281
			j->anchorA = or_rotate( qA.c, qA.s, j->localOriginAnchorA.x - lcA.x, j->localOriginAnchorA.y - lcA.y );
			j->anchorB = or_rotate( qB.c, qB.s, j->localOriginAnchorB.x - lcB.x, j->localOriginAnchorB.y - lcB.y );
			j->axisA = or_rotate( qA.c, qA.s, j->localAxisA.x, j->localAxisA.y );
			j->deltaCenter = ( orVec2r ){ (real)simB->center.x - (real)simA->center.x, (real)simB->center.y - (real)simA->center.y };
			j->deltaAngle = or_relative_angle( qB.c, qB.s, qA.c, qA.s ) - (real)j->referenceAngle;
			orVec2r rA = j->anchorA, rB = j->anchorB;
			orVec2r d = or_add( j->deltaCenter, or_sub( rB, rA ) );
			real a1 = or_cross( or_add( d, rA ), j->axisA ), a2 = or_cross( rB, j->axisA );
			real k = mA + mB + iA * a1 * a1 + iB * a2 * a2;
			j->axialMass = k > 0 ? (real)1 / k : 0;
			j->springSoftness = or_make_softness( j->hertz, j->dampingRatio, h );
			if ( ctx->enableWarmStarting == false )
			{
				j->linearImpulse = ( orVec2r ){ 0, 0 };
				j->springImpulse = 0;
				j->motorImpulse = 0;
				j->lowerImpulse = 0;
				j->upperImpulse = 0;
			}
			break;
		}
		case b2_wheelJoint:
		{
			// src/wheel_joint.c:222-267
			j->anchorA = or_rotate( qA.c, qA.s, j->localOriginAnchorA.x - lcA.x, j->localOriginAnchorA.y - lcA.y );
			j->anchorB = or_rotate( qB.c, qB.s, j->localOriginAnchorB.x - lcB.x, j->localOriginAnchorB.y - lcB.y );
			j->axisA = or_rotate( qA.c, qA.s, j->localAxisA.x, j->localAxisA.y );
			j->deltaCenter = ( orVec2r ){ (real)simB->center.x - (real)simA->center.x, (real)simB->center.y - (real)simA->center.y };
			orVec2r rA = j->anchorA, rB = j->anchorB;
			orVec2r d = or_add( j->deltaCenter, or_sub( rB, rA ) );
			orVec2r axisA = j->axisA, perpA = or_left_perp( axisA );
			real s1 = or_cross( or_add( d, rA ), perpA ), s2 = or_cross( rB, perpA );
			real kp = mA + mB + iA * s1 * s1 + iB * s2 * s2;
			j->perpMass = kp > 0 ? (real)1 / kp : 0;
			real a1 = or_cross( or_add( d, rA ), axisA ), a2 = or_cross( rB, axisA );
			real ka = mA + mB + iA * a1 * a1 + iB * a2 * a2;
			j->axialMass = ka > 0 ? (real)1 / ka : 0;
			j->springSoftness = or_make_softness( j->hertz, j->dampingRatio, h );
			real km = iA + iB;
			j->motorMass = km > 0 ? (real)1 / km : 0;
			if ( ctx->enableWarmStarting == false )
			{
				j->impulse = 0;
				j->springImpulse = 0;
				j->motorImpulse = 0;
				j->lowerImpulse = 0;
				j->upperImpulse = 0;
			}
			break;
		}
		case b2_motorJoint:
		{
			// src/motor_joint.c:137-166
			j->anchorA = or_rotate( qA.c, qA.s, j->localOriginAnchorA.x - lcA.x, j->localOriginAnchorA.y - lcA.y );
			j->anchorB = or_rotate( qB.c, qB.s, j->localOriginAnchorB.x - lcB.x, j->localOriginAnchorB.y - lcB.y );
			// b2Sub( b2Sub( centerB, centerA ), linearOffset )
			real dcx = (real)simB->center.x - (real)simA->center.x, dcy = (real)simB->center.y - (real)simA->center.y;
			j->deltaCenter = ( orVec2r ){ dcx - (real)j->linearOffset.x, dcy - (real)j->linearOffset.y };
			j->deltaAngle = or_relative_angle( qB.c, qB.s, qA.c, qA.s ) - (real)j->angularOffset;
			j->deltaAngle = or_unwind( j->deltaAngle );
			orVec2r rA = j->anchorA, rB = j->anchorB;
			real kxx = mA + mB + rA.y * rA.y * iA + rB.y * rB.y * iB;
			real kxy = -rA.y * rA.x * iA - rB.y * rB.x * iB;
			real kyy = mA + mB + rA.x * rA.x * iA + rB.x * rB.x * iB;
			// K = { cx = (kxx, kxy), cy = (kxy, kyy) }: a = kxx, b = cy.x = kxy, c = cx.y = kxy, d = kyy
			j->linearMass = or_inverse22( kxx, kxy, kxy, kyy );
			real ka = iA + iB;
			j->angularMass = ka > 0 ? (real)1 / ka : 0;
			if ( ctx->enableWarmStarting == false )
			{
				j->linearImpulse = ( orVec2r ){ 0, 0 };
				j->angularImpulse = 0;
			}
			break;
		}
		case b2_mouseJoint:
		{
			// src/mouse_joint.c:100-125 (body B only)
			j->anchorB = or_rotate( qB.c, qB.s, j->localOriginAnchorB.x - lcB.x, j->localOriginAnchorB.y - lcB.y );
			j->linearSoftness = or_make_softness( j->hertz, j->dampingRatio, h );
			j->angularSoftness = or_make_softness( 0.5f, 0.1f, h );
			orVec2r rB = j->anchorB;
			real kxx = mB + iB * rB.y * rB.y;
			real kxy = -iB * rB.x * rB.y;
			real kyy = mB + iB * rB.x * rB.x;
			j->linearMass = or_inverse22( kxx, kxy, kxy, kyy );
			j->deltaCenter = ( orVec2r ){ (real)simB->center.x - (real)j->target.x, (real)simB->center.y - (real)j->target.y };
			if ( ctx->enableWarmStarting == false )
			{
				j->linearImpulse = ( orVec2r ){ 0, 0 };
				j->angularImpulse = 0;
			}
			break;
		}
		default:
			OR_ASSERT( false );
	}
}

// b2WarmStart*Joint (src/revolute_joint.c:271-297, src/weld_joint.c:174-198)
void or_warm_start_joint( orJointSim* j, orBodyState* states )
{
	orBodyState dA, dB;
	orBodyState* stateA = or_joint_state( states, j->indexA, &dA );
	orBodyState* stateB = or_joint_state( states, j->indexB, &dB );
	real mA = j->invMassA, mB = j->invMassB, iA = j->invIA, iB = j->invIB;
	orVec2r rA = or_rotate( stateA->dq.x, stateA->dq.y, j->anchorA.x, j->anchorA.y );
	orVec2r rB = or_rotate( stateB->dq.x, stateB->dq.y, j->anchorB.x, j->anchorB.y );
	real axialImpulse;
	switch ( j->type )
	{
		case b2_revoluteJoint:
			axialImpulse = j->springImpulse + j->motorImpulse + j->lowerImpulse - j->upperImpulse;
			break;
		case b2_weldJoint:
		case b2_motorJoint: // src/motor_joint.c:170-193
			axialImpulse = j->angularImpulse;
			break;
		case b2_distanceJoint:
		{
			// src/distance_joint.c:262-285
			orVec2r ds = or_add( or_sub( or_dp( stateB ), or_dp( stateA ) ), or_sub( rB, rA ) );
			orVec2r axis = or_normalize( or_add( j->deltaCenter, ds ) );
			real axial = j->impulse + j->lowerImpulse - j->upperImpulse + j->motorImpulse;
			orVec2r P = or_scale( axial, axis );
			or_apply( &stateA->v, &stateA->w, &stateB->v, &stateB->w, j, P, or_cross( rA, P ), or_cross( rB, P ) );
			return;
		}
		case b2_prismaticJoint:
		case b2_wheelJoint:
		{
			// src/prismatic_joint.c:284-321, src/wheel_joint.c:270-307
			orVec2r d = or_add( or_add( or_sub( or_dp( stateB ), or_dp( stateA ) ), j->deltaCenter ), or_sub( rB, rA ) );
			orVec2r axisA = or_rotate( stateA->dq.x, stateA->dq.y, j->axisA.x, j->axisA.y );
			orVec2r perpA = or_left_perp( axisA );
			real a1 = or_cross( or_add( d, rA ), axisA ), a2 = or_cross( rB, axisA );
			real s1 = or_cross( or_add( d, rA ), perpA ), s2 = or_cross( rB, perpA );
			real axial, perp, angular;
			if ( j->type == b2_prismaticJoint )
			{
				axial = j->springImpulse + j->motorImpulse + j->lowerImpulse - j->upperImpulse;
				perp = j->linearImpulse.x;
				angular = j->linearImpulse.y;
			}
			else
			{
				axial = j->springImpulse + j->lowerImpulse - j->upperImpulse;
				perp = j->impulse;
				angular = j->motorImpulse;
			}
			orVec2r P = or_add( or_scale( axial, axisA ), or_scale( perp, perpA ) );
			real LA = axial * a1 + perp * s1 + angular;
			real LB = axial * a2 + perp * s2 + angular;
			or_apply( &stateA->v, &stateA->w, &stateB->v, &stateB->w, j, P, LA, LB );
			return;
		}
		case b2_mouseJoint:
		{
			// src/mouse_joint.c:128-147: body B only
			orVec2r vB = stateB->v;
			real wB = stateB->w;
			vB.x = vB.x + mB * j->linearImpulse.x;
			vB.y = vB.y + mB * j->linearImpulse.y;
			wB += iB * ( or_cross( rB, j->linearImpulse ) + j->angularImpulse );
			stateB->v = vB;
			stateB->w = wB;
			return;
		}
		default:
			OR_ASSERT( false );
			return;
	}
	orVec2r P = j->linearImpulse;
	stateA->v.x = stateA->v.x - mA * P.x;
	stateA->v.y = stateA->v.y - mA * P.y;
	stateA->w -= iA * ( ( rA.x * P.y - rA.y * P.x ) + axialImpulse );
	stateB->v.x = stateB->v.x + mB * P.x;
	stateB->v.y = stateB->v.y + mB * P.y;
	stateB->w += iB * ( ( rB.x * P.y - rB.y * P.x ) + axialImpulse );
}

// the point-to-point block of the revolute and weld solvers (src/revolute_joint.c:422-458)
static void or_point_constraint( orJointSim* j, orBodyState* stateA, orBodyState* stateB, orVec2r* vA, real* wA, orVec2r* vB, real* wB,
								 orSoftness soft, bool applyBias )
{
	real mA = j->invMassA, mB = j->invMassB, iA = j->invIA, iB = j->invIB;
	orVec2r rA = or_rotate( stateA->dq.x, stateA->dq.y, j->anchorA.x, j->anchorA.y );
	orVec2r rB = or_rotate( stateB->dq.x, stateB->dq.y, j->anchorB.x, j->anchorB.y );
	// Cdot = (vB + cross(wB, rB)) - (vA + cross(wA, rA)), cross(w, r) = (-w r.y, w r.x)
	real cdx = ( vB->x + ( -*wB ) * rB.y ) - ( vA->x + ( -*wA ) * rA.y );
	real cdy = ( vB->y + *wB * rB.x ) - ( vA->y + *wA * rA.x );
	real bx = 0, by = 0, massScale = 1, impulseScale = 0;
	if ( applyBias )
	{
		// separation = ((dcB - dcA) + (rB - rA)) + deltaCenter
		real sx = ( ( stateB->dp.x - stateA->dp.x ) + ( rB.x - rA.x ) ) + j->deltaCenter.x;
		real sy = ( ( stateB->dp.y - stateA->dp.y ) + ( rB.y - rA.y ) ) + j->deltaCenter.y;
		bx = soft.biasRate * sx;
		by = soft.biasRate * sy;
		massScale = soft.massScale;
		impulseScale = soft.impulseScale;
	}
	real k11 = mA + mB + rA.y * rA.y * iA + rB.y * rB.y * iB;
	real k12 = -rA.y * rA.x * iA - rB.y * rB.x * iB;
	real k22 = mA + mB + rA.x * rA.x * iA + rB.x * rB.x * iB;
	orVec2r b = or_solve22( k11, k12, k12, k22, cdx + bx, cdy + by );
	real px = -massScale * b.x - impulseScale * j->linearImpulse.x;
	real py = -massScale * b.y - impulseScale * j->linearImpulse.y;
	j->linearImpulse.x += px;
	j->linearImpulse.y += py;
	vA->x = vA->x - mA * px;
	vA->y = vA->y - mA * py;
	*wA -= iA * ( rA.x * py - rA.y * px );
	vB->x = vB->x + mB * px;
	vB->y = vB->y + mB * py;
	*wB += iB * ( rB.x * py - rB.y * px );
}

// src/revolute_joint.c:299-467
static void or_solve_revolute( orJointSim* j, orBodyState* stateA, orBodyState* stateB, const orStepContext* ctx, bool useBias )
{
	real iA = j->invIA, iB = j->invIB;
	orVec2r vA = stateA->v, vB = stateB->v;
	real wA = stateA->w, wB = stateB->w;
	bool fixedRotation = ( iA + iB == 0 );
	orSoftness js = ctx->jointSoftness;
	if ( j->enableSpring && fixedRotation == false )
	{
		real C = or_relative_angle( stateB->dq.x, stateB->dq.y, stateA->dq.x, stateA->dq.y ) + j->deltaAngle;
		real bias = j->springSoftness.biasRate * C;
		real Cdot = wB - wA;
		real impulse = -j->springSoftness.massScale * j->axialMass * ( Cdot + bias ) - j->springSoftness.impulseScale * j->springImpulse;
		j->springImpulse += impulse;
		wA -= iA * impulse;
		wB += iB * impulse;
	}
	if ( j->enableMotor && fixedRotation == false )
	{
		real Cdot = wB - wA - j->motorSpeed;
		real impulse = -j->axialMass * Cdot;
		real oldImpulse = j->motorImpulse;
		real maxImpulse = ctx->h * j->maxMotor;
		j->motorImpulse = or_clampr( j->motorImpulse + impulse, -maxImpulse, maxImpulse );
		impulse = j->motorImpulse - oldImpulse;
		wA -= iA * impulse;
		wB += iB * impulse;
	}
	if ( j->enableLimit && fixedRotation == false )
	{
		real jointAngle = or_relative_angle( stateB->dq.x, stateB->dq.y, stateA->dq.x, stateA->dq.y ) + j->deltaAngle;
		jointAngle = or_unwind( jointAngle );
		{
			real C = jointAngle - j->lower;
			real bias = 0, massScale = 1, impulseScale = 0;
			if ( C > 0 )
				bias = C * ctx->inv_h;
			else if ( useBias )
			{
				bias = js.biasRate * C;
				massScale = js.massScale;
				impulseScale = js.impulseScale;
			}
			real Cdot = wB - wA;
			real impulse = -massScale * j->axialMass * ( Cdot + bias ) - impulseScale * j->lowerImpulse;
			real oldImpulse = j->lowerImpulse;
			j->lowerImpulse = or_maxr( j->lowerImpulse + impulse, 0 );
			impulse = j->lowerImpulse - oldImpulse;
			wA -= iA * impulse;
			wB += iB * impulse;
		}
		{
			// upper limit: the reference scales the lower impulse (src/revolute_joint.c:405)
			real C = j->upper - jointAngle;
			real bias = 0, massScale = 1, impulseScale = 0;
			if ( C > 0 )
				bias = C * ctx->inv_h;
			else if ( useBias )
			{
				bias = js.biasRate * C;
				massScale = js.massScale;
				impulseScale = js.impulseScale;
			}
			real Cdot = wA - wB;
			real impulse = -massScale * j->axialMass * ( Cdot + bias ) - impulseScale * j->lowerImpulse;
			real oldImpulse = j->upperImpulse;
			j->upperImpulse = or_maxr( j->upperImpulse + impulse, 0 );
			impulse = j->upperImpulse - oldImpulse;
			wA += iA * impulse;
			wB -= iB * impulse;
		}
	}
	or_point_constraint( j, stateA, stateB, &vA, &wA, &vB, &wB, js, useBias );
	stateA->v = vA;
	stateA->w = wA;
	stateB->v = vB;
	stateB->w = wB;
}

// src/weld_joint.c:200-297
static void or_solve_weld( orJointSim* j, orBodyState* stateA, orBodyState* stateB, bool useBias )
{
	real iA = j->invIA, iB = j->invIB;
	orVec2r vA = stateA->v, vB = stateB->v;
	real wA = stateA->w, wB = stateB->w;
	{
		real bias = 0, massScale = 1, impulseScale = 0;
		if ( useBias || j->angularHertz > 0.0f )
		{
			real C = or_relative_angle( stateB->dq.x, stateB->dq.y, stateA->dq.x, stateA->dq.y ) + j->deltaAngle;
			bias = j->angularSoftness.biasRate * C;
			massScale = j->angularSoftness.massScale;
			impulseScale = j->angularSoftness.impulseScale;
		}
		real Cdot = wB - wA;
		real impulse = -massScale * j->axialMass * ( Cdot + bias ) - impulseScale * j->angularImpulse;
		j->angularImpulse += impulse;
		wA -= iA * impulse;
		wB += iB * impulse;
	}
	or_point_constraint( j, stateA, stateB, &vA, &wA, &vB, &wB, j->linearSoftness, useBias || j->linearHertz > 0.0f );
	stateA->v = vA;
	stateA->w = wA;
	stateB->v = vB;
	stateB->w = wB;
}

// src/distance_joint.c:288-461
static void or_solve_distance( orJointSim* j, orBodyState* stateA, orBodyState* stateB, const orStepContext* ctx, bool useBias )
{
	orVec2r vA = stateA->v, vB = stateB->v;
	real wA = stateA->w, wB = stateB->w;
	orSoftness js = ctx->jointSoftness;
	orVec2r rA = or_rotate( stateA->dq.x, stateA->dq.y, j->anchorA.x, j->anchorA.y );
	orVec2r rB = or_rotate( stateB->dq.x, stateB->dq.y, j->anchorB.x, j->anchorB.y );
	orVec2r ds = or_add( or_sub( or_dp( stateB ), or_dp( stateA ) ), or_sub( rB, rA ) );
	orVec2r separation = or_add( j->deltaCenter, ds );
	real length = or_length( separation );
	orVec2r axis = or_normalize( separation );
	// (vB - vA) + (wB x rB - wA x rA)
#define OR_VR() or_add( or_sub( vB, vA ), or_sub( or_cross_sv( wB, rB ), or_cross_sv( wA, rA ) ) )
	if ( j->enableSpring && ( j->lower < j->upper || j->enableLimit == false ) )
	{
		if ( j->hertz > 0.0f )
		{
			real Cdot = or_dot( axis, OR_VR() );
			real C = length - j->length;
			real bias = j->springSoftness.biasRate * C;
			real m = j->springSoftness.massScale * j->axialMass;
			real impulse = -m * ( Cdot + bias ) - j->springSoftness.impulseScale * j->impulse;
			j->impulse += impulse;
			orVec2r P = or_scale( impulse, axis );
			or_apply( &vA, &wA, &vB, &wB, j, P, or_cross( rA, P ), or_cross( rB, P ) );
		}
		if ( j->enableLimit )
		{
			{
				real Cdot = or_dot( axis, OR_VR() );
				real C = length - j->lower;
				real bias = 0, massCoeff = 1, impulseCoeff = 0;
				if ( C > 0 )
					bias = C * ctx->inv_h;
				else if ( useBias )
				{
					bias = js.biasRate * C;
					massCoeff = js.massScale;
					impulseCoeff = js.impulseScale;
				}
				real impulse = -massCoeff * j->axialMass * ( Cdot + bias ) - impulseCoeff * j->lowerImpulse;
				real newImpulse = or_maxr( 0, j->lowerImpulse + impulse );
				impulse = newImpulse - j->lowerImpulse;
				j->lowerImpulse = newImpulse;
				orVec2r P = or_scale( impulse, axis );
				or_apply( &vA, &wA, &vB, &wB, j, P, or_cross( rA, P ), or_cross( rB, P ) );
			}
			{
				// (vA - vB) + (wA x rA - wB x rB)
				orVec2r vr = or_add( or_sub( vA, vB ), or_sub( or_cross_sv( wA, rA ), or_cross_sv( wB, rB ) ) );
				real Cdot = or_dot( axis, vr );
				real C = j->upper - length;
				real bias = 0, massScale = 1, impulseScale = 0;
				if ( C > 0 )
					bias = C * ctx->inv_h;
				else if ( useBias )
				{
					bias = js.biasRate * C;
					massScale = js.massScale;
					impulseScale = js.impulseScale;
				}
				real impulse = -massScale * j->axialMass * ( Cdot + bias ) - impulseScale * j->upperImpulse;
				real newImpulse = or_maxr( 0, j->upperImpulse + impulse );
				impulse = newImpulse - j->upperImpulse;
				j->upperImpulse = newImpulse;
				orVec2r P = or_scale( -impulse, axis );
				or_apply( &vA, &wA, &vB, &wB, j, P, or_cross( rA, P ), or_cross( rB, P ) );
			}
		}
		if ( j->enableMotor )
		{
			real Cdot = or_dot( axis, OR_VR() );
			real impulse = j->axialMass * ( j->motorSpeed - Cdot );
			real oldImpulse = j->motorImpulse;
			real maxImpulse = ctx->h * j->maxMotor;
			j->motorImpulse = or_clampr( j->motorImpulse + impulse, -maxImpulse, maxImpulse );
			impulse = j->motorImpulse - oldImpulse;
			orVec2r P = or_scale( impulse, axis );
			or_apply( &vA, &wA, &vB, &wB, j, P, or_cross( rA, P ), or_cross( rB, P ) );
		}
	}
	else
	{
		real Cdot = or_dot( axis, OR_VR() );
		real C = length - j->length;
		real bias = 0, massScale = 1, impulseScale = 0;
		if ( useBias )
		{
			bias = js.biasRate * C;
			massScale = js.massScale;
			impulseScale = js.impulseScale;
		}
		real impulse = -massScale * j->axialMass * ( Cdot + bias ) - impulseScale * j->impulse;
		j->impulse += impulse;
		orVec2r P = or_scale( impulse, axis );
		or_apply( &vA, &wA, &vB, &wB, j, P, or_cross( rA, P ), or_cross( rB, P ) );
	}
#undef OR_VR
	stateA->v = vA;
	stateA->w = wA;
	stateB->v = vB;
	stateB->w = wB;
}

// the axial constraints shared by the prismatic and wheel joints' lower and upper limits
// (src/prismatic_joint.c:410-473, src/wheel_joint.c:369-435)
static void or_axis_limits( orJointSim* j, orVec2r axisA, real a1, real a2, real translation, orVec2r* vA, real* wA, orVec2r* vB,
							real* wB, const orStepContext* ctx, bool useBias, bool prismatic )
{
	orSoftness js = ctx->jointSoftness;
	{
		real C = translation - j->lower;
		real bias = 0, massScale = 1, impulseScale = 0;
		if ( C > 0 )
			bias = C * ctx->inv_h;
		else if ( useBias )
		{
			bias = js.biasRate * C;
			massScale = js.massScale;
			impulseScale = js.impulseScale;
		}
		real oldImpulse = j->lowerImpulse;
		real Cdot = or_dot( axisA, or_sub( *vB, *vA ) ) + a2 * *wB - a1 * *wA;
		// the prismatic joint writes -axialMass * massScale * (...), the wheel -massScale * axialMass
		real impulse = prismatic ? -j->axialMass * massScale * ( Cdot + bias ) - impulseScale * oldImpulse
								 : -massScale * j->axialMass * ( Cdot + bias ) - impulseScale * oldImpulse;
		j->lowerImpulse = or_maxr( oldImpulse + impulse, 0 );
		impulse = j->lowerImpulse - oldImpulse;
		or_apply( vA, wA, vB, wB, j, or_scale( impulse, axisA ), impulse * a1, impulse * a2 );
	}
	{
		real C = j->upper - translation;
		real bias = 0, massScale = 1, impulseScale = 0;
		if ( C > 0 )
			bias = C * ctx->inv_h;
		else if ( useBias )
		{
			bias = js.biasRate * C;
			massScale = js.massScale;
			impulseScale = js.impulseScale;
		}
		real oldImpulse = j->upperImpulse;
		real Cdot = or_dot( axisA, or_sub( *vA, *vB ) ) + a1 * *wA - a2 * *wB;
		real impulse = prismatic ? -j->axialMass * massScale * ( Cdot + bias ) - impulseScale * oldImpulse
								 : -massScale * j->axialMass * ( Cdot + bias ) - impulseScale * oldImpulse;
		j->upperImpulse = or_maxr( oldImpulse + impulse, 0 );
		impulse = j->upperImpulse - oldImpulse;
		orVec2r P = or_scale( impulse, axisA );
		real LA = impulse * a1, LB = impulse * a2;
		vA->x = vA->x + j->invMassA * P.x;
		vA->y = vA->y + j->invMassA * P.y;
		*wA += j->invIA * LA;
		vB->x = vB->x - j->invMassB * P.x;
		vB->y = vB->y - j->invMassB * P.y;
		*wB -= j->invIB * LB;
	}
}

// src/prismatic_joint.c:323-480
static void or_solve_prismatic( orJointSim* j, orBodyState* stateA, orBodyState* stateB, const orStepContext* ctx, bool useBias )
{
	orVec2r vA = stateA->v, vB = stateB->v;
	real wA = stateA->w, wB = stateB->w;
	orSoftness js = ctx->jointSoftness;
	real mA = j->invMassA, mB = j->invMassB, iA = j->invIA, iB = j->invIB;
	orVec2r rA = or_rotate( stateA->dq.x, stateA->dq.y, j->anchorA.x, j->anchorA.y );
	orVec2r rB = or_rotate( stateB->dq.x, stateB->dq.y, j->anchorB.x, j->anchorB.y );
	orVec2r d = or_add( or_add( or_sub( or_dp( stateB ), or_dp( stateA ) ), j->deltaCenter ), or_sub( rB, rA ) );
	orVec2r axisA = or_rotate( stateA->dq.x, stateA->dq.y, j->axisA.x, j->axisA.y );
	real translation = or_dot( axisA, d );
	real a1 = or_cross( or_add( d, rA ), axisA ), a2 = or_cross( rB, axisA );
	if ( j->enableSpring )
	{
		real C = translation;
		real bias = j->springSoftness.biasRate * C;
		real Cdot = or_dot( axisA, or_sub( vB, vA ) ) + a2 * wB - a1 * wA;
		real impulse = -j->springSoftness.massScale * j->axialMass * ( Cdot + bias ) - j->springSoftness.impulseScale * j->springImpulse;
		j->springImpulse += impulse;
		or_apply( &vA, &wA, &vB, &wB, j, or_scale( impulse, axisA ), impulse * a1, impulse * a2 );
	}
	if ( j->enableMotor )
	{
		real Cdot = or_dot( axisA, or_sub( vB, vA ) ) + a2 * wB - a1 * wA;
		real impulse = j->axialMass * ( j->motorSpeed - Cdot );
		real oldImpulse = j->motorImpulse;
		real maxImpulse = ctx->h * j->maxMotor;
		j->motorImpulse = or_clampr( j->motorImpulse + impulse, -maxImpulse, maxImpulse );
		impulse = j->motorImpulse - oldImpulse;
		or_apply( &vA, &wA, &vB, &wB, j, or_scale( impulse, axisA ), impulse * a1, impulse * a2 );
	}
	if ( j->enableLimit )
		or_axis_limits( j, axisA, a1, a2, translation, &vA, &wA, &vB, &wB, ctx, useBias, true );
	{
		orVec2r perpA = or_left_perp( axisA );
		real s1 = or_cross( or_add( d, rA ), perpA ), s2 = or_cross( rB, perpA );
		real cdx = or_dot( perpA, or_sub( vB, vA ) ) + s2 * wB - s1 * wA;
		real cdy = wB - wA;
		real bx = 0, by = 0, massScale = 1, impulseScale = 0;
		if ( useBias )
		{
			real Cx = or_dot( perpA, d );
			real Cy = or_relative_angle( stateB->dq.x, stateB->dq.y, stateA->dq.x, stateA->dq.y ) + j->deltaAngle;
			bx = js.biasRate * Cx;
			by = js.biasRate * Cy;
			massScale = js.massScale;
			impulseScale = js.impulseScale;
		}
		real k11 = mA + mB + iA * s1 * s1 + iB * s2 * s2;
		real k12 = iA * s1 + iB * s2;
		real k22 = iA + iB;
		if ( k22 == 0 )
			k22 = 1;
		orVec2r b = or_solve22( k11, k12, k12, k22, cdx + bx, cdy + by );
		real ix = -massScale * b.x - impulseScale * j->linearImpulse.x;
		real iy = -massScale * b.y - impulseScale * j->linearImpulse.y;
		j->linearImpulse.x += ix;
		j->linearImpulse.y += iy;
		orVec2r P = or_scale( ix, perpA );
		real LA = ix * s1 + iy, LB = ix * s2 + iy;
		or_apply( &vA, &wA, &vB, &wB, j, P, LA, LB );
	}
	stateA->v = vA;
	stateA->w = wA;
	stateB->v = vB;
	stateB->w = wB;
}

// src/wheel_joint.c:309-464
static void or_solve_wheel( orJointSim* j, orBodyState* stateA, orBodyState* stateB, const orStepContext* ctx, bool useBias )
{
	orVec2r vA = stateA->v, vB = stateB->v;
	real wA = stateA->w, wB = stateB->w;
	orSoftness js = ctx->jointSoftness;
	real iA = j->invIA, iB = j->invIB;
	bool fixedRotation = ( iA + iB == 0 );
	orVec2r rA = or_rotate( stateA->dq.x, stateA->dq.y, j->anchorA.x, j->anchorA.y );
	orVec2r rB = or_rotate( stateB->dq.x, stateB->dq.y, j->anchorB.x, j->anchorB.y );
	orVec2r d = or_add( or_add( or_sub( or_dp( stateB ), or_dp( stateA ) ), j->deltaCenter ), or_sub( rB, rA ) );
	orVec2r axisA = or_rotate( stateA->dq.x, stateA->dq.y, j->axisA.x, j->axisA.y );
	real translation = or_dot( axisA, d );
	real a1 = or_cross( or_add( d, rA ), axisA ), a2 = or_cross( rB, axisA );
	if ( j->enableMotor && fixedRotation == false )
	{
		real Cdot = wB - wA - j->motorSpeed;
		real impulse = -j->motorMass * Cdot;
		real oldImpulse = j->motorImpulse;
		real maxImpulse = ctx->h * j->maxMotor;
		j->motorImpulse = or_clampr( j->motorImpulse + impulse, -maxImpulse, maxImpulse );
		impulse = j->motorImpulse - oldImpulse;
		wA -= iA * impulse;
		wB += iB * impulse;
	}
	if ( j->enableSpring )
	{
		real C = translation;
		real bias = j->springSoftness.biasRate * C;
		real Cdot = or_dot( axisA, or_sub( vB, vA ) ) + a2 * wB - a1 * wA;
		real impulse = -j->springSoftness.massScale * j->axialMass * ( Cdot + bias ) - j->springSoftness.impulseScale * j->springImpulse;
		j->springImpulse += impulse;
		or_apply( &vA, &wA, &vB, &wB, j, or_scale( impulse, axisA ), impulse * a1, impulse * a2 );
	}
	if ( j->enableLimit )
		or_axis_limits( j, axisA, a1, a2, or_dot( axisA, d ), &vA, &wA, &vB, &wB, ctx, useBias, false );
	{
		orVec2r perpA = or_left_perp( axisA );
		real bias = 0, massScale = 1, impulseScale = 0;
		if ( useBias )
		{
			real C = or_dot( perpA, d );
			bias = js.biasRate * C;
			massScale = js.massScale;
			impulseScale = js.impulseScale;
		}
		real s1 = or_cross( or_add( d, rA ), perpA ), s2 = or_cross( rB, perpA );
		real Cdot = or_dot( perpA, or_sub( vB, vA ) ) + s2 * wB - s1 * wA;
		real impulse = -massScale * j->perpMass * ( Cdot + bias ) - impulseScale * j->impulse;
		j->impulse += impulse;
		or_apply( &vA, &wA, &vB, &wB, j, or_scale( impulse, perpA ), impulse * s1, impulse * s2 );
	}
	stateA->v = vA;
	stateA->w = wA;
	stateB->v = vB;
	stateB->w = wB;
}

// src/motor_joint.c:195-270
static void or_solve_motor( orJointSim* j, orBodyState* stateA, orBodyState* stateB, const orStepContext* ctx )
{
	orVec2r vA = stateA->v, vB = stateB->v;
	real wA = stateA->w, wB = stateB->w;
	real mA = j->invMassA, mB = j->invMassB, iA = j->invIA, iB = j->invIB;
	{
		real angularSeperation = or_relative_angle( stateB->dq.x, stateB->dq.y, stateA->dq.x, stateA->dq.y ) + j->deltaAngle;
		angularSeperation = or_unwind( angularSeperation );
		real angularBias = (real)ctx->inv_h * (real)j->correctionFactor * angularSeperation;
		real Cdot = wB - wA;
		real impulse = -j->angularMass * ( Cdot + angularBias );
		real oldImpulse = j->angularImpulse;
		real maxImpulse = ctx->h * j->maxTorque;
		j->angularImpulse = or_clampr( j->angularImpulse + impulse, -maxImpulse, maxImpulse );
		impulse = j->angularImpulse - oldImpulse;
		wA -= iA * impulse;
		wB += iB * impulse;
	}
	{
		orVec2r rA = or_rotate( stateA->dq.x, stateA->dq.y, j->anchorA.x, j->anchorA.y );
		orVec2r rB = or_rotate( stateB->dq.x, stateB->dq.y, j->anchorB.x, j->anchorB.y );
		orVec2r ds = or_add( or_sub( or_dp( stateB ), or_dp( stateA ) ), or_sub( rB, rA ) );
		orVec2r linearSeparation = or_add( j->deltaCenter, ds );
		orVec2r linearBias = or_scale( (real)ctx->inv_h * (real)j->correctionFactor, linearSeparation );
		orVec2r Cdot = or_sub( or_add( vB, or_cross_sv( wB, rB ) ), or_add( vA, or_cross_sv( wA, rA ) ) );
		orVec2r b = or_mul_mv( j->linearMass, or_add( Cdot, linearBias ) );
		orVec2r impulse = { -b.x, -b.y };
		orVec2r oldImpulse = j->linearImpulse;
		real maxImpulse = ctx->h * j->maxForce;
		j->linearImpulse = or_add( j->linearImpulse, impulse );
		if ( or_dot( j->linearImpulse, j->linearImpulse ) > maxImpulse * maxImpulse )
		{
			j->linearImpulse = or_normalize( j->linearImpulse );
			j->linearImpulse.x *= maxImpulse;
			j->linearImpulse.y *= maxImpulse;
		}
		impulse = or_sub( j->linearImpulse, oldImpulse );
		vA.x = vA.x - mA * impulse.x;
		vA.y = vA.y - mA * impulse.y;
		wA -= iA * or_cross( rA, impulse );
		vB.x = vB.x + mB * impulse.x;
		vB.y = vB.y + mB * impulse.y;
		wB += iB * or_cross( rB, impulse );
	}
	stateA->v = vA;
	stateA->w = wA;
	stateB->v = vB;
	stateB->w = wB;
}

// src/mouse_joint.c:149-212 (body B only)
static void or_solve_mouse( orJointSim* j, orBodyState* stateB, const orStepContext* ctx )
{
	real mB = j->invMassB, iB = j->invIB;
	orVec2r vB = stateB->v;
	real wB = stateB->w;
	{
		real massScale = j->angularSoftness.massScale;
		real impulseScale = j->angularSoftness.impulseScale;
		real impulse = iB > 0 ? -wB / iB : 0;
		impulse = massScale * impulse - impulseScale * j->angularImpulse;
		j->angularImpulse += impulse;
		wB += iB * impulse;
	}
	real maxImpulse = j->maxForce * ctx->h;
	{
		orVec2r rB = or_rotate( stateB->dq.x, stateB->dq.y, j->anchorB.x, j->anchorB.y );
		orVec2r Cdot = or_add( vB, or_cross_sv( wB, rB ) );
		orVec2r separation = or_add( or_add( or_dp( stateB ), rB ), j->deltaCenter );
		orVec2r bias = or_scale( j->linearSoftness.biasRate, separation );
		real massScale = j->linearSoftness.massScale;
		real impulseScale = j->linearSoftness.impulseScale;
		orVec2r b = or_mul_mv( j->linearMass, or_add( Cdot, bias ) );
		orVec2r impulse;
		impulse.x = -massScale * b.x - impulseScale * j->linearImpulse.x;
		impulse.y = -massScale * b.y - impulseScale * j->linearImpulse.y;
		orVec2r oldImpulse = j->linearImpulse;
		j->linearImpulse.x += impulse.x;
		j->linearImpulse.y += impulse.y;
		real mag = or_length( j->linearImpulse );
		if ( mag > maxImpulse )
			j->linearImpulse = or_scale( maxImpulse, or_normalize( j->linearImpulse ) );
		impulse.x = j->linearImpulse.x - oldImpulse.x;
		impulse.y = j->linearImpulse.y - oldImpulse.y;
		vB.x = vB.x + mB * impulse.x;
		vB.y = vB.y + mB * impulse.y;
		wB += iB * or_cross( rB, impulse );
	}
	stateB->v = vB;
	stateB->w = wB;
}

// b2SolveJoint (src/joint.c:1068-1104)
void or_solve_joint( orJointSim* j, orBodyState* states, const orStepContext* ctx, bool useBias )
{
	orBodyState dA, dB;
	orBodyState* stateA = or_joint_state( states, j->indexA, &dA );
	orBodyState* stateB = or_joint_state( states, j->indexB, &dB );
	switch ( j->type )
	{
		case b2_revoluteJoint:
			or_solve_revolute( j, stateA, stateB, ctx, useBias );
			break;
		case b2_weldJoint:
			or_solve_weld( j, stateA, stateB, useBias );
			break;
		case b2_distanceJoint:
			or_solve_distance( j, stateA, stateB, ctx, useBias );
			break;
		case b2_prismaticJoint:
			or_solve_prismatic( j, stateA, stateB, ctx, useBias );
			break;
		case b2_wheelJoint:
			or_solve_wheel( j, stateA, stateB, ctx, useBias );
			break;
		case b2_motorJoint:
			or_solve_motor( j, stateA, stateB, ctx );
			break;
		case b2_mouseJoint:
			or_solve_mouse( j, stateB, ctx );
			break;
		default:
			OR_ASSERT( false );
	}
}
