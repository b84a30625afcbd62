// Joint API and the host half of the joints: ids, the bodies' joint lists, graph colouring beside
// the contacts, the pair filter joints add, and the joints' device records with their host mirror.
// The step kernels prepare, warm start and solve the joints (joints.hpp). Line references are to
// the reference.
#include "world_internal.hpp"

#include <cmath>

namespace b2mi
{

void destroyContactImpl( World* w, Contact* c );
void b2mi_internal_beginEdit( World* w );

// ---- device records ----

// brings the host mirror of the joint records up to date (a step has moved the impulses on)
static void ensureJointMirror( World* w )
{
	const int n = w->jointIdPool.nextIndex;
	if ( (int)w->devJoints.size() < n )
		w->devJoints.resize( n );
	if ( w->jointsDirty || w->jointsEpoch == stepEpoch( w ) || w->dJoints == nullptr )
	{
		w->jointsEpoch = stepEpoch( w );
		return;
	}
	const int onDevice = std::min( n, w->capJoints );
	if ( onDevice > 0 )
	{
		HIP_CHECK( hipMemcpyAsync( w->devJoints.data(), w->dJoints, sizeof( DevJoint ) * onDevice, hipMemcpyDeviceToHost, stream() ) );
		HIP_CHECK( hipStreamSynchronize( stream() ) );
	}
	w->jointsEpoch = stepEpoch( w );
}

// writes the joint records to the device before a step that follows joint edits
void uploadJoints( World* w )
{
	const int n = w->jointIdPool.nextIndex;
	B2MI_CHECK( (int)w->devJoints.size() >= n );
	if ( w->capJoints < n || w->dJoints == nullptr )
	{
		if ( w->dJoints )
		{
			HIP_CHECK( hipStreamSynchronize( stream() ) );
			HIP_CHECK( hipFree( w->dJoints ) );
		}
		w->capJoints = std::max( 16, n + n / 2 );
		HIP_CHECK( hipMalloc( (void**)&w->dJoints, sizeof( DevJoint ) * (size_t)w->capJoints ) );
	}
	if ( n > 0 )
	{
		HIP_CHECK( hipMemcpyAsync( w->dJoints, w->devJoints.data(), sizeof( DevJoint ) * n, hipMemcpyHostToDevice, stream() ) );
		HIP_CHECK( hipStreamSynchronize( stream() ) );
	}
	w->jointsDirty = false;
	w->jointsEpoch = stepEpoch( w );
}

static DevJoint& editJoint( World* w, const Joint* j )
{
	B2MI_CHECK( worldLocked( w ) == false );
	ensureJointMirror( w );
	w->jointsDirty = true;
	markActive( w );
	return w->devJoints[j->jointId];
}

static const DevJoint& readJoint( World* w, const Joint* j )
{
	ensureJointMirror( w );
	return w->devJoints[j->jointId];
}

// 1/h of the world's last step (src/world.c:731), 0 before its first
static float lastInvH( const World* w )
{
	const StepGroup* g = w->group;
	if ( g == nullptr || g->epoch == 0 || w->groupIndex < 0 || w->groupIndex >= (int)g->hWorlds.size() )
		return 0.0f;
	return g->hWorlds[w->groupIndex].inv_h;
}

// ---- colouring (src/constraint_graph.c:216-324) ----

static int assignJointColor( World* w, int bodyIdA, int bodyIdB, bool staticA, bool staticB )
{
	B2MI_CHECK( staticA == false || staticB == false );
	for ( int i = 0; i < kOverflow; ++i )
	{
		BitSet& set = w->colorBodySets[i];
		if ( !staticA && !staticB )
		{
			if ( set.get( bodyIdA ) || set.get( bodyIdB ) )
				continue;
			set.set( bodyIdA );
			set.set( bodyIdB );
			return i;
		}
		if ( !staticA )
		{
			if ( set.get( bodyIdA ) )
				continue;
			set.set( bodyIdA );
			return i;
		}
		if ( set.get( bodyIdB ) )
			continue;
		set.set( bodyIdB );
		return i;
	}
	return kOverflow;
}

static void removeJointFromList( World* w, std::vector<int>& list, int localIndex )
{
	const int last = (int)list.size() - 1;
	if ( localIndex < last )
	{
		list[localIndex] = list[last];
		w->joints[list[localIndex]].localIndex = localIndex;
	}
	list.pop_back();
}

// ---- lookup ----

static Body* jointBody( World* w, b2BodyId id )
{
	const int index = id.index1 - 1;
	B2MI_CHECK( 0 <= index && index < (int)w->bodies.size() );
	Body* b = &w->bodies[index];
	B2MI_CHECK( b->id == index && b->revision == id.revision );
	return b;
}

static Joint* getJoint( World* w, b2JointId id )
{
	const int index = id.index1 - 1;
	B2MI_CHECK( 0 <= index && index < (int)w->joints.size() );
	Joint* j = &w->joints[index];
	B2MI_CHECK( j->jointId == index && j->revision == id.revision );
	return j;
}

#define B2MI_JOINT( jointId )               \
	World* w = worldAt( ( jointId ).world0 ); \
	Joint* joint = getJoint( w, ( jointId ) )

#define B2MI_JOINT_OF( jointId, kind )   \
	B2MI_JOINT( jointId );               \
	B2MI_CHECK( joint->type == ( kind ) )

static b2BodyId bodyIdOf( const World* w, int bodyId )
{
	return b2BodyId{ bodyId + 1, (uint16_t)w->slot, w->bodies[bodyId].revision };
}

// ---- creation and destruction (src/joint.c:130-318, 660-756) ----

static Joint* createJoint( World* w, Body* bodyA, Body* bodyB, void* userData, float drawSize, b2JointType type, bool collideConnected,
						   b2JointId* outId )
{
	B2MI_CHECK( worldLocked( w ) == false );
	B2MI_CHECK( bodyA->setIndex >= 0 && bodyB->setIndex >= 0 );
	// the reference joins no island for a joint of two non-dynamic awake bodies (src/island.c:315)
	if ( !( bodyA->setIndex == kStaticSet && bodyB->setIndex == kStaticSet ) && bodyA->type != b2_dynamicBody &&
		 bodyB->type != b2_dynamicBody )
		fatal( "a joint needs a dynamic body (the reference links it into no island otherwise)", __FILE__, __LINE__ );
	ensureJointMirror( w );
	const int jointId = w->jointIdPool.alloc();
	if ( jointId == (int)w->joints.size() )
		w->joints.push_back( Joint{} );
	if ( (int)w->devJoints.size() <= jointId )
		w->devJoints.resize( jointId + 1 );
	Joint& j = w->joints[jointId];
	j.jointId = jointId;
	j.userData = userData;
	j.revision += 1;
	j.drawSize = drawSize;
	j.type = type;
	j.collideConnected = collideConnected;
	Body* bodies[2] = { bodyA, bodyB };
	for ( int e = 0; e < 2; ++e )
	{
		Body* b = bodies[e];
		j.edges[e].bodyId = b->id;
		j.edges[e].prevKey = -1;
		j.edges[e].nextKey = b->headJointKey;
		const int key = ( jointId << 1 ) | e;
		if ( b->headJointKey != -1 )
			w->joints[b->headJointKey >> 1].edges[b->headJointKey & 1].prevKey = key;
		b->headJointKey = key;
		b->jointCount += 1;
	}
	if ( bodyA->setIndex == kStaticSet && bodyB->setIndex == kStaticSet )
	{
		j.setIndex = kStaticSet;
		j.colorIndex = -1;
		j.localIndex = (int)w->staticJoints.size();
		w->staticJoints.push_back( jointId );
	}
	else
	{
		j.setIndex = kAwakeSet;
		j.colorIndex = assignJointColor( w, bodyA->id, bodyB->id, bodyA->setIndex == kStaticSet, bodyB->setIndex == kStaticSet );
		j.localIndex = (int)w->jointColors[j.colorIndex].size();
		w->jointColors[j.colorIndex].push_back( jointId );
	}
	DevJoint& d = w->devJoints[jointId];
	memset( &d, 0, sizeof( d ) );
	d.type = type;
	w->jointsDirty = true;
	w->topologyDirty = true;
	markActive( w );
	*outId = b2JointId{ jointId + 1, (uint16_t)w->slot, j.revision };
	return &j;
}

// src/joint.c:283-318 (the bodies are not woken)
static void destroyContactsBetweenBodies( World* w, Body* bodyA, Body* bodyB )
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
	while ( contactKey != -1 )
	{
		Contact* c = &w->contacts[contactKey >> 1];
		const int edgeIndex = contactKey & 1;
		contactKey = c->edges[edgeIndex].nextKey;
		if ( c->edges[edgeIndex ^ 1].bodyId == otherBodyId )
			destroyContactImpl( w, c );
	}
}

void destroyJointImpl( World* w, Joint* j )
{
	const int jointId = j->jointId;
	for ( int e = 0; e < 2; ++e )
	{
		JointEdge& edge = j->edges[e];
		Body& b = w->bodies[edge.bodyId];
		if ( edge.prevKey != -1 )
			w->joints[edge.prevKey >> 1].edges[edge.prevKey & 1].nextKey = edge.nextKey;
		if ( edge.nextKey != -1 )
			w->joints[edge.nextKey >> 1].edges[edge.nextKey & 1].prevKey = edge.prevKey;
		const int key = ( jointId << 1 ) | e;
		if ( b.headJointKey == key )
			b.headJointKey = edge.nextKey;
		b.jointCount -= 1;
	}
	if ( j->setIndex == kAwakeSet )
	{
		if ( j->colorIndex != kOverflow )
		{
			w->colorBodySets[j->colorIndex].clear( j->edges[0].bodyId );
			w->colorBodySets[j->colorIndex].clear( j->edges[1].bodyId );
		}
		removeJointFromList( w, w->jointColors[j->colorIndex], j->localIndex );
	}
	else
		removeJointFromList( w, w->staticJoints, j->localIndex );
	j->setIndex = -1;
	j->colorIndex = -1;
	j->localIndex = -1;
	j->jointId = -1;
	w->jointIdPool.free( jointId );
	w->topologyDirty = true;
	markActive( w );
}

// src/body.c:1724-1760: whether a joint between two bodies turns their collision off
bool jointsAllowCollision( const World* w, const Body& bodyA, const Body& bodyB )
{
	int jointKey, otherBodyId;
	if ( bodyA.jointCount < bodyB.jointCount )
	{
		jointKey = bodyA.headJointKey;
		otherBodyId = bodyB.id;
	}
	else
	{
		jointKey = bodyB.headJointKey;
		otherBodyId = bodyA.id;
	}
	while ( jointKey != -1 )
	{
		const Joint& j = w->joints[jointKey >> 1];
		const int edgeIndex = jointKey & 1;
		if ( j.collideConnected == false && j.edges[edgeIndex ^ 1].bodyId == otherBodyId )
			return false;
		jointKey = j.edges[edgeIndex].nextKey;
	}
	return true;
}

// the joints of a body about to be destroyed go first (src/body.c:349-362)
void destroyBodyJoints( World* w, Body* body )
{
	for ( int key = body->headJointKey; key != -1; )
	{
		Joint* j = &w->joints[key >> 1];
		key = j->edges[key & 1].nextKey;
		destroyJointImpl( w, j );
	}
}

// colour counts of the world's awake joints (b2World_GetCounters adds them to the contacts')
int jointColorCount( const World* w, int color )
{
	return (int)w->jointColors[color].size();
}

// b2RelativeAngle then b2UnwindAngle on the host (include/box2d/math_functions.h:387-409)
static float relativeAngle( b2Rot b, b2Rot a )
{
	float s = b.s * a.c - b.c * a.s;
	float c = b.c * a.c + b.s * a.s;
	return atan2f( s, c );
}

static float unwindAngle( float angle )
{
	if ( angle < -b2_pi )
		return angle + 2.0f * b2_pi;
	if ( angle > b2_pi )
		return angle - 2.0f * b2_pi;
	return angle;
}

static void setFlag( DevJoint& d, uint32_t flag, bool on )
{
	d.flags = on ? ( d.flags | flag ) : ( d.flags & ~flag );
}

} // namespace b2mi

using namespace b2mi;

#define B2MI_JOINT_DEF( def )                         \
	World* w = getWorld( worldId );                   \
	B2MI_CHECK( ( def ) != nullptr && ( def )->internalValue == kCookie ); \
	Body* bodyA = jointBody( w, ( def )->bodyIdA );   \
	Body* bodyB = jointBody( w, ( def )->bodyIdB );   \
	b2JointId jointId

extern "C" {

// ---- definitions (src/joint.c:19-80) ----

B2_API b2DistanceJointDef b2DefaultDistanceJointDef( void )
{
	b2DistanceJointDef def{};
	def.length = 1.0f;
	def.maxLength = kHuge;
	def.internalValue = kCookie;
	return def;
}

B2_API b2MotorJointDef b2DefaultMotorJointDef( void )
{
	b2MotorJointDef def{};
	def.maxForce = 1.0f;
	def.maxTorque = 1.0f;
	def.correctionFactor = 0.3f;
	def.internalValue = kCookie;
	return def;
}

B2_API b2MouseJointDef b2DefaultMouseJointDef( void )
{
	b2MouseJointDef def{};
	def.hertz = 4.0f;
	def.dampingRatio = 1.0f;
	def.maxForce = 1.0f;
	def.internalValue = kCookie;
	return def;
}

B2_API b2PrismaticJointDef b2DefaultPrismaticJointDef( void )
{
	b2PrismaticJointDef def{};
	def.localAxisA = b2Vec2{ 1.0f, 0.0f };
	def.internalValue = kCookie;
	return def;
}

B2_API b2RevoluteJointDef b2DefaultRevoluteJointDef( void )
{
	b2RevoluteJointDef def{};
	def.drawSize = 0.25f;
	def.internalValue = kCookie;
	return def;
}

B2_API b2WeldJointDef b2DefaultWeldJointDef( void )
{
	b2WeldJointDef def{};
	def.internalValue = kCookie;
	return def;
}

B2_API b2WheelJointDef b2DefaultWheelJointDef( void )
{
	b2WheelJointDef def{};
	def.localAxisA.y = 1.0f;
	def.enableSpring = true;
	def.hertz = 1.0f;
	def.dampingRatio = 0.7f;
	def.internalValue = kCookie;
	return def;
}

// ---- creation (src/joint.c:320-658) ----

B2_API b2JointId b2CreateRevoluteJoint( b2WorldId worldId, const b2RevoluteJointDef* def )
{
	B2MI_JOINT_DEF( def );
	createJoint( w, bodyA, bodyB, def->userData, def->drawSize, b2_revoluteJoint, def->collideConnected, &jointId );
	DevJoint& d = w->devJoints[jointId.index1 - 1];
	d.localOriginAnchorA = def->localAnchorA;
	d.localOriginAnchorB = def->localAnchorB;
	DevRevoluteJoint& r = d.revolute;
	r.referenceAngle = b2ClampFloat( def->referenceAngle, -b2_pi, b2_pi );
	r.hertz = def->hertz;
	r.dampingRatio = def->dampingRatio;
	r.lowerAngle = b2MinFloat( def->lowerAngle, def->upperAngle );
	r.upperAngle = b2MaxFloat( def->lowerAngle, def->upperAngle );
	r.lowerAngle = b2ClampFloat( r.lowerAngle, -b2_pi, b2_pi );
	r.upperAngle = b2ClampFloat( r.upperAngle, -b2_pi, b2_pi );
	r.maxMotorTorque = def->maxMotorTorque;
	r.motorSpeed = def->motorSpeed;
	setFlag( d, kJointSpring, def->enableSpring );
	setFlag( d, kJointLimit, def->enableLimit );
	setFlag( d, kJointMotor, def->enableMotor );
	if ( def->collideConnected == false )
		destroyContactsBetweenBodies( w, bodyA, bodyB );
	return jointId;
}

B2_API b2JointId b2CreateWeldJoint( b2WorldId worldId, const b2WeldJointDef* def )
{
	B2MI_JOINT_DEF( def );
	createJoint( w, bodyA, bodyB, def->userData, 1.0f, b2_weldJoint, def->collideConnected, &jointId );
	DevJoint& d = w->devJoints[jointId.index1 - 1];
	d.localOriginAnchorA = def->localAnchorA;
	d.localOriginAnchorB = def->localAnchorB;
	DevWeldJoint& j = d.weld;
	j.referenceAngle = def->referenceAngle;
	j.linearHertz = def->linearHertz;
	j.linearDampingRatio = def->linearDampingRatio;
	j.angularHertz = def->angularHertz;
	j.angularDampingRatio = def->angularDampingRatio;
	if ( def->collideConnected == false )
		destroyContactsBetweenBodies( w, bodyA, bodyB );
	return jointId;
}

B2_API b2JointId b2CreateDistanceJoint( b2WorldId worldId, const b2DistanceJointDef* def )
{
	B2MI_JOINT_DEF( def );
	B2MI_CHECK( b2IsValid( def->length ) && def->length > 0.0f );
	createJoint( w, bodyA, bodyB, def->userData, 1.0f, b2_distanceJoint, def->collideConnected, &jointId );
	DevJoint& d = w->devJoints[jointId.index1 - 1];
	d.localOriginAnchorA = def->localAnchorA;
	d.localOriginAnchorB = def->localAnchorB;
	DevDistanceJoint& j = d.distance;
	j.length = b2MaxFloat( def->length, kLinearSlop );
	j.hertz = def->hertz;
	j.dampingRatio = def->dampingRatio;
	j.minLength = b2MaxFloat( def->minLength, kLinearSlop );
	j.maxLength = b2MaxFloat( def->minLength, def->maxLength );
	j.maxMotorForce = def->maxMotorForce;
	j.motorSpeed = def->motorSpeed;
	setFlag( d, kJointSpring, def->enableSpring );
	setFlag( d, kJointLimit, def->enableLimit );
	setFlag( d, kJointMotor, def->enableMotor );
	if ( def->collideConnected == false )
		destroyContactsBetweenBodies( w, bodyA, bodyB );
	return jointId;
}

B2_API b2JointId b2CreateMotorJoint( b2WorldId worldId, const b2MotorJointDef* def )
{
	B2MI_JOINT_DEF( def );
	createJoint( w, bodyA, bodyB, def->userData, 1.0f, b2_motorJoint, def->collideConnected, &jointId );
	DevJoint& d = w->devJoints[jointId.index1 - 1];
	d.localOriginAnchorA = b2Vec2{ 0.0f, 0.0f };
	d.localOriginAnchorB = b2Vec2{ 0.0f, 0.0f };
	DevMotorJoint& j = d.motor;
	j.linearOffset = def->linearOffset;
	j.angularOffset = def->angularOffset;
	j.maxForce = def->maxForce;
	j.maxTorque = def->maxTorque;
	j.correctionFactor = b2ClampFloat( def->correctionFactor, 0.0f, 1.0f );
	if ( def->collideConnected == false )
		destroyContactsBetweenBodies( w, bodyA, bodyB );
	return jointId;
}

// src/joint.c:407-439: the anchors are the target in each body's frame; contacts stay
B2_API b2JointId b2CreateMouseJoint( b2WorldId worldId, const b2MouseJointDef* def )
{
	B2MI_JOINT_DEF( def );
	b2Transform transformA = b2Body_GetTransform( def->bodyIdA );
	b2Transform transformB = b2Body_GetTransform( def->bodyIdB );
	createJoint( w, bodyA, bodyB, def->userData, 1.0f, b2_mouseJoint, def->collideConnected, &jointId );
	DevJoint& d = w->devJoints[jointId.index1 - 1];
	d.localOriginAnchorA = b2InvTransformPoint( transformA, def->target );
	d.localOriginAnchorB = b2InvTransformPoint( transformB, def->target );
	DevMouseJoint& j = d.mouse;
	j.targetA = def->target;
	j.hertz = def->hertz;
	j.dampingRatio = def->dampingRatio;
	j.maxForce = def->maxForce;
	return jointId;
}

B2_API b2JointId b2CreatePrismaticJoint( b2WorldId worldId, const b2PrismaticJointDef* def )
{
	B2MI_JOINT_DEF( def );
	createJoint( w, bodyA, bodyB, def->userData, 1.0f, b2_prismaticJoint, def->collideConnected, &jointId );
	DevJoint& d = w->devJoints[jointId.index1 - 1];
	d.localOriginAnchorA = def->localAnchorA;
	d.localOriginAnchorB = def->localAnchorB;
	DevPrismaticJoint& j = d.prismatic;
	j.localAxisA = b2Normalize( def->localAxisA );
	j.referenceAngle = def->referenceAngle;
	j.hertz = def->hertz;
	j.dampingRatio = def->dampingRatio;
	j.lowerTranslation = def->lowerTranslation;
	j.upperTranslation = def->upperTranslation;
	j.maxMotorForce = def->maxMotorForce;
	j.motorSpeed = def->motorSpeed;
	setFlag( d, kJointSpring, def->enableSpring );
	setFlag( d, kJointLimit, def->enableLimit );
	setFlag( d, kJointMotor, def->enableMotor );
	if ( def->collideConnected == false )
		destroyContactsBetweenBodies( w, bodyA, bodyB );
	return jointId;
}

B2_API b2JointId b2CreateWheelJoint( b2WorldId worldId, const b2WheelJointDef* def )
{
	B2MI_JOINT_DEF( def );
	createJoint( w, bodyA, bodyB, def->userData, 1.0f, b2_wheelJoint, def->collideConnected, &jointId );
	DevJoint& d = w->devJoints[jointId.index1 - 1];
	d.localOriginAnchorA = def->localAnchorA;
	d.localOriginAnchorB = def->localAnchorB;
	DevWheelJoint& j = d.wheel;
	j.localAxisA = b2Normalize( def->localAxisA );
	j.lowerTranslation = def->lowerTranslation;
	j.upperTranslation = def->upperTranslation;
	j.maxMotorTorque = def->maxMotorTorque;
	j.motorSpeed = def->motorSpeed;
	j.hertz = def->hertz;
	j.dampingRatio = def->dampingRatio;
	setFlag( d, kJointSpring, def->enableSpring );
	setFlag( d, kJointLimit, def->enableLimit );
	setFlag( d, kJointMotor, def->enableMotor );
	if ( def->collideConnected == false )
		destroyContactsBetweenBodies( w, bodyA, bodyB );
	return jointId;
}

B2_API void b2DestroyJoint( b2JointId jointId )
{
	B2MI_JOINT( jointId );
	B2MI_CHECK( worldLocked( w ) == false );
	ensureJointMirror( w );
	destroyJointImpl( w, joint );
}

// ---- generic queries (src/joint.c:758-981) ----

B2_API bool b2Joint_IsValid( b2JointId id )
{
	World* w = b2mi::worldAt( id.world0 );
	if ( w == nullptr )
		return false;
	const int index = id.index1 - 1;
	if ( index < 0 || index >= (int)w->joints.size() )
		return false;
	const Joint& j = w->joints[index];
	return j.jointId == index && j.revision == id.revision;
}

B2_API b2JointType b2Joint_GetType( b2JointId jointId )
{
	B2MI_JOINT( jointId );
	return joint->type;
}

B2_API b2BodyId b2Joint_GetBodyA( b2JointId jointId )
{
	B2MI_JOINT( jointId );
	return bodyIdOf( w, joint->edges[0].bodyId );
}

B2_API b2BodyId b2Joint_GetBodyB( b2JointId jointId )
{
	B2MI_JOINT( jointId );
	return bodyIdOf( w, joint->edges[1].bodyId );
}

B2_API b2Vec2 b2Joint_GetLocalAnchorA( b2JointId jointId )
{
	B2MI_JOINT( jointId );
	return readJoint( w, joint ).localOriginAnchorA;
}

B2_API b2Vec2 b2Joint_GetLocalAnchorB( b2JointId jointId )
{
	B2MI_JOINT( jointId );
	return readJoint( w, joint ).localOriginAnchorB;
}

// src/joint.c:806-846: turning collision on makes the smaller body's proxies look for pairs again;
// turning it off destroys the contacts between the bodies
B2_API void b2Joint_SetCollideConnected( b2JointId jointId, bool shouldCollide )
{
	B2MI_JOINT( jointId );
	B2MI_CHECK( worldLocked( w ) == false );
	if ( joint->collideConnected == shouldCollide )
		return;
	joint->collideConnected = shouldCollide;
	Body* bodyA = &w->bodies[joint->edges[0].bodyId];
	Body* bodyB = &w->bodies[joint->edges[1].bodyId];
	markActive( w );
	if ( shouldCollide )
	{
		int shapeId = bodyA->shapeCount < bodyB->shapeCount ? bodyA->headShapeId : bodyB->headShapeId;
		while ( shapeId != -1 )
		{
			const Shape& sh = w->shapes[shapeId];
			if ( sh.proxyKey != -1 && !w->moveSet.get( (uint32_t)sh.proxyKey ) )
			{
				w->moveSet.set( (uint32_t)sh.proxyKey );
				w->moveArray.push_back( sh.proxyKey );
			}
			shapeId = sh.nextShapeId;
		}
	}
	else
		destroyContactsBetweenBodies( w, bodyA, bodyB );
}

B2_API bool b2Joint_GetCollideConnected( b2JointId jointId )
{
	B2MI_JOINT( jointId );
	return joint->collideConnected;
}

B2_API void b2Joint_SetUserData( b2JointId jointId, void* userData )
{
	B2MI_JOINT( jointId );
	joint->userData = userData;
}

B2_API void* b2Joint_GetUserData( b2JointId jointId )
{
	B2MI_JOINT( jointId );
	return joint->userData;
}

// every body stays awake here (no sleep), so there is nothing to wake
B2_API void b2Joint_WakeBodies( b2JointId jointId )
{
	B2MI_JOINT( jointId );
	(void)w;
}

B2_API b2Vec2 b2Joint_GetConstraintForce( b2JointId jointId )
{
	B2MI_JOINT( jointId );
	const DevJoint& d = readJoint( w, joint );
	const float inv_h = lastInvH( w );
	switch ( joint->type )
	{
		case b2_revoluteJoint: // src/revolute_joint.c:158-162
			return b2MulSV( inv_h, d.revolute.linearImpulse );
		case b2_weldJoint: // src/weld_joint.c:66-70
			return b2MulSV( inv_h, d.weld.linearImpulse );
		case b2_motorJoint: // src/motor_joint.c:74-78
			return b2MulSV( inv_h, d.motor.linearImpulse );
		case b2_mouseJoint: // src/mouse_joint.c:66-70
			return b2MulSV( inv_h, d.mouse.linearImpulse );
		case b2_distanceJoint:
		{
			// src/distance_joint.c:183-196: along the current anchor axis
			b2Transform xfA = b2Body_GetTransform( bodyIdOf( w, joint->edges[0].bodyId ) );
			b2Transform xfB = b2Body_GetTransform( bodyIdOf( w, joint->edges[1].bodyId ) );
			b2Vec2 pA = b2TransformPoint( xfA, d.localOriginAnchorA );
			b2Vec2 pB = b2TransformPoint( xfB, d.localOriginAnchorB );
			b2Vec2 axis = b2Normalize( b2Sub( pB, pA ) );
			const DevDistanceJoint& j = d.distance;
			float force = ( j.impulse + j.lowerImpulse - j.upperImpulse + j.motorImpulse ) * inv_h;
			return b2MulSV( force, axis );
		}
		case b2_prismaticJoint:
		{
			// src/prismatic_joint.c:144-160
			b2Transform xfA = b2Body_GetTransform( bodyIdOf( w, joint->edges[0].bodyId ) );
			const DevPrismaticJoint& j = d.prismatic;
			b2Vec2 axisA = b2RotateVector( xfA.q, j.localAxisA );
			b2Vec2 perpA = b2LeftPerp( axisA );
			float perpForce = inv_h * j.impulse.x;
			float axialForce = inv_h * ( j.motorImpulse + j.lowerImpulse - j.upperImpulse );
			return b2Add( b2MulSV( perpForce, perpA ), b2MulSV( axialForce, axisA ) );
		}
		case b2_wheelJoint:
		{
			// src/wheel_joint.c:145-158: the axis of the last prepare
			const DevWheelJoint& j = d.wheel;
			b2Vec2 perpA = b2LeftPerp( j.axisA );
			float perpForce = inv_h * j.perpImpulse;
			float axialForce = inv_h * ( j.springImpulse + j.lowerImpulse - j.upperImpulse );
			return b2Add( b2MulSV( perpForce, perpA ), b2MulSV( axialForce, j.axisA ) );
		}
		default:
			fatal( "unknown joint type", __FILE__, __LINE__ );
			return b2Vec2{ 0.0f, 0.0f };
	}
}

B2_API float b2Joint_GetConstraintTorque( b2JointId jointId )
{
	B2MI_JOINT( jointId );
	const DevJoint& d = readJoint( w, joint );
	const float inv_h = lastInvH( w );
	switch ( joint->type )
	{
		case b2_revoluteJoint: // src/revolute_joint.c:164-169
			return inv_h * ( d.revolute.motorImpulse + d.revolute.lowerImpulse - d.revolute.upperImpulse );
		case b2_weldJoint: // src/weld_joint.c:72-75
			return inv_h * d.weld.angularImpulse;
		case b2_motorJoint: // src/motor_joint.c:80-83
			return inv_h * d.motor.angularImpulse;
		case b2_mouseJoint: // src/mouse_joint.c:72-75
			return inv_h * d.mouse.angularImpulse;
		case b2_distanceJoint:
			return 0.0f;
		case b2_prismaticJoint: // src/prismatic_joint.c:162-165
			return inv_h * d.prismatic.impulse.y;
		case b2_wheelJoint: // src/wheel_joint.c:160-163
			return inv_h * d.wheel.motorImpulse;
		default:
			fatal( "unknown joint type", __FILE__, __LINE__ );
			return 0.0f;
	}
}

// ---- the joints of a body (src/body.c:1700-1722) ----

B2_API int b2Body_GetJointCount( b2BodyId bodyId )
{
	World* w = worldAt( bodyId.world0 );
	return jointBody( w, bodyId )->jointCount;
}

B2_API int b2Body_GetJoints( b2BodyId bodyId, b2JointId* jointArray, int capacity )
{
	World* w = worldAt( bodyId.world0 );
	Body* body = jointBody( w, bodyId );
	int jointKey = body->headJointKey, index = 0;
	while ( jointKey != -1 && index < capacity )
	{
		const Joint& j = w->joints[jointKey >> 1];
		jointArray[index++] = b2JointId{ j.jointId + 1, (uint16_t)w->slot, j.revision };
		jointKey = j.edges[jointKey & 1].nextKey;
	}
	return index;
}

// ---- revolute joint (src/revolute_joint.c:18-156) ----

B2_API void b2RevoluteJoint_EnableSpring( b2JointId jointId, bool enableSpring )
{
	B2MI_JOINT_OF( jointId, b2_revoluteJoint );
	if ( enableSpring != ( ( readJoint( w, joint ).flags & kJointSpring ) != 0 ) )
	{
		DevJoint& d = editJoint( w, joint );
		setFlag( d, kJointSpring, enableSpring );
		d.revolute.springImpulse = 0.0f;
	}
}

B2_API void b2RevoluteJoint_SetSpringHertz( b2JointId jointId, float hertz )
{
	B2MI_JOINT_OF( jointId, b2_revoluteJoint );
	editJoint( w, joint ).revolute.hertz = hertz;
}

B2_API float b2RevoluteJoint_GetSpringHertz( b2JointId jointId )
{
	B2MI_JOINT_OF( jointId, b2_revoluteJoint );
	return readJoint( w, joint ).revolute.hertz;
}

B2_API void b2RevoluteJoint_SetSpringDampingRatio( b2JointId jointId, float dampingRatio )
{
	B2MI_JOINT_OF( jointId, b2_revoluteJoint );
	editJoint( w, joint ).revolute.dampingRatio = dampingRatio;
}

B2_API float b2RevoluteJoint_GetSpringDampingRatio( b2JointId jointId )
{
	B2MI_JOINT_OF( jointId, b2_revoluteJoint );
	return readJoint( w, joint ).revolute.dampingRatio;
}

B2_API float b2RevoluteJoint_GetAngle( b2JointId jointId )
{
	B2MI_JOINT_OF( jointId, b2_revoluteJoint );
	b2Transform transformA = b2Body_GetTransform( bodyIdOf( w, joint->edges[0].bodyId ) );
	b2Transform transformB = b2Body_GetTransform( bodyIdOf( w, joint->edges[1].bodyId ) );
	float angle = relativeAngle( transformB.q, transformA.q ) - readJoint( w, joint ).revolute.referenceAngle;
	return unwindAngle( angle );
}

B2_API void b2RevoluteJoint_EnableLimit( b2JointId jointId, bool enableLimit )
{
	B2MI_JOINT_OF( jointId, b2_revoluteJoint );
	if ( enableLimit != ( ( readJoint( w, joint ).flags & kJointLimit ) != 0 ) )
	{
		DevJoint& d = editJoint( w, joint );
		setFlag( d, kJointLimit, enableLimit );
		d.revolute.lowerImpulse = 0.0f;
		d.revolute.upperImpulse = 0.0f;
	}
}

B2_API bool b2RevoluteJoint_IsLimitEnabled( b2JointId jointId )
{
	B2MI_JOINT_OF( jointId, b2_revoluteJoint );
	return ( readJoint( w, joint ).flags & kJointLimit ) != 0;
}

B2_API float b2RevoluteJoint_GetLowerLimit( b2JointId jointId )
{
	B2MI_JOINT_OF( jointId, b2_revoluteJoint );
	return readJoint( w, joint ).revolute.lowerAngle;
}

B2_API float b2RevoluteJoint_GetUpperLimit( b2JointId jointId )
{
	B2MI_JOINT_OF( jointId, b2_revoluteJoint );
	return readJoint( w, joint ).revolute.upperAngle;
}

B2_API void b2RevoluteJoint_SetLimits( b2JointId jointId, float lower, float upper )
{
	B2MI_JOINT_OF( jointId, b2_revoluteJoint );
	const DevRevoluteJoint& r = readJoint( w, joint ).revolute;
	if ( lower != r.lowerAngle || upper != r.upperAngle )
	{
		DevRevoluteJoint& e = editJoint( w, joint ).revolute;
		e.lowerAngle = b2MinFloat( lower, upper );
		e.upperAngle = b2MaxFloat( lower, upper );
		e.lowerImpulse = 0.0f;
		e.upperImpulse = 0.0f;
	}
}

B2_API void b2RevoluteJoint_EnableMotor( b2JointId jointId, bool enableMotor )
{
	B2MI_JOINT_OF( jointId, b2_revoluteJoint );
	if ( enableMotor != ( ( readJoint( w, joint ).flags & kJointMotor ) != 0 ) )
	{
		DevJoint& d = editJoint( w, joint );
		setFlag( d, kJointMotor, enableMotor );
		d.revolute.motorImpulse = 0.0f;
	}
}

B2_API bool b2RevoluteJoint_IsMotorEnabled( b2JointId jointId )
{
	B2MI_JOINT_OF( jointId, b2_revoluteJoint );
	return ( readJoint( w, joint ).flags & kJointMotor ) != 0;
}

B2_API void b2RevoluteJoint_SetMotorSpeed( b2JointId jointId, float motorSpeed )
{
	B2MI_JOINT_OF( jointId, b2_revoluteJoint );
	editJoint( w, joint ).revolute.motorSpeed = motorSpeed;
}

B2_API float b2RevoluteJoint_GetMotorSpeed( b2JointId jointId )
{
	B2MI_JOINT_OF( jointId, b2_revoluteJoint );
	return readJoint( w, joint ).revolute.motorSpeed;
}

B2_API float b2RevoluteJoint_GetMotorTorque( b2JointId jointId )
{
	B2MI_JOINT_OF( jointId, b2_revoluteJoint );
	return lastInvH( w ) * readJoint( w, joint ).revolute.motorImpulse;
}

B2_API void b2RevoluteJoint_SetMaxMotorTorque( b2JointId jointId, float torque )
{
	B2MI_JOINT_OF( jointId, b2_revoluteJoint );
	editJoint( w, joint ).revolute.maxMotorTorque = torque;
}

B2_API float b2RevoluteJoint_GetMaxMotorTorque( b2JointId jointId )
{
	B2MI_JOINT_OF( jointId, b2_revoluteJoint );
	return readJoint( w, joint ).revolute.maxMotorTorque;
}

// ---- weld joint (src/weld_joint.c:17-64) ----

B2_API void b2WeldJoint_SetLinearHertz( b2JointId jointId, float hertz )
{
	B2MI_JOINT_OF( jointId, b2_weldJoint );
	editJoint( w, joint ).weld.linearHertz = hertz;
}

B2_API float b2WeldJoint_GetLinearHertz( b2JointId jointId )
{
	B2MI_JOINT_OF( jointId, b2_weldJoint );
	return readJoint( w, joint ).weld.linearHertz;
}

B2_API void b2WeldJoint_SetLinearDampingRatio( b2JointId jointId, float dampingRatio )
{
	B2MI_JOINT_OF( jointId, b2_weldJoint );
	editJoint( w, joint ).weld.linearDampingRatio = dampingRatio;
}

B2_API float b2WeldJoint_GetLinearDampingRatio( b2JointId jointId )
{
	B2MI_JOINT_OF( jointId, b2_weldJoint );
	return readJoint( w, joint ).weld.linearDampingRatio;
}

B2_API void b2WeldJoint_SetAngularHertz( b2JointId jointId, float hertz )
{
	B2MI_JOINT_OF( jointId, b2_weldJoint );
	editJoint( w, joint ).weld.angularHertz = hertz;
}

B2_API float b2WeldJoint_GetAngularHertz( b2JointId jointId )
{
	B2MI_JOINT_OF( jointId, b2_weldJoint );
	return readJoint( w, joint ).weld.angularHertz;
}

B2_API void b2WeldJoint_SetAngularDampingRatio( b2JointId jointId, float dampingRatio )
{
	B2MI_JOINT_OF( jointId, b2_weldJoint );
	editJoint( w, joint ).weld.angularDampingRatio = dampingRatio;
}

B2_API float b2WeldJoint_GetAngularDampingRatio( b2JointId jointId )
{
	B2MI_JOINT_OF( jointId, b2_weldJoint );
	return readJoint( w, joint ).weld.angularDampingRatio;
}

// ---- distance joint (src/distance_joint.c:17-181) ----

B2_API void b2DistanceJoint_SetLength( b2JointId jointId, float length )
{
	B2MI_JOINT_OF( jointId, b2_distanceJoint );
	DevDistanceJoint& j = editJoint( w, joint ).distance;
	j.length = b2ClampFloat( length, kLinearSlop, kHuge );
	j.impulse = 0.0f;
	j.lowerImpulse = 0.0f;
	j.upperImpulse = 0.0f;
}

B2_API float b2DistanceJoint_GetLength( b2JointId jointId )
{
	B2MI_JOINT_OF( jointId, b2_distanceJoint );
	return readJoint( w, joint ).distance.length;
}

B2_API void b2DistanceJoint_EnableSpring( b2JointId jointId, bool enableSpring )
{
	B2MI_JOINT_OF( jointId, b2_distanceJoint );
	setFlag( editJoint( w, joint ), kJointSpring, enableSpring );
}

B2_API bool b2DistanceJoint_IsSpringEnabled( b2JointId jointId )
{
	B2MI_JOINT_OF( jointId, b2_distanceJoint );
	return ( readJoint( w, joint ).flags & kJointSpring ) != 0;
}

B2_API void b2DistanceJoint_SetSpringHertz( b2JointId jointId, float hertz )
{
	B2MI_JOINT_OF( jointId, b2_distanceJoint );
	editJoint( w, joint ).distance.hertz = hertz;
}

B2_API void b2DistanceJoint_SetSpringDampingRatio( b2JointId jointId, float dampingRatio )
{
	B2MI_JOINT_OF( jointId, b2_distanceJoint );
	editJoint( w, joint ).distance.dampingRatio = dampingRatio;
}

B2_API float b2DistanceJoint_GetHertz( b2JointId jointId )
{
	B2MI_JOINT_OF( jointId, b2_distanceJoint );
	return readJoint( w, joint ).distance.hertz;
}

B2_API float b2DistanceJoint_GetDampingRatio( b2JointId jointId )
{
	B2MI_JOINT_OF( jointId, b2_distanceJoint );
	return readJoint( w, joint ).distance.dampingRatio;
}

B2_API void b2DistanceJoint_EnableLimit( b2JointId jointId, bool enableLimit )
{
	B2MI_JOINT_OF( jointId, b2_distanceJoint );
	setFlag( editJoint( w, joint ), kJointLimit, enableLimit );
}

B2_API bool b2DistanceJoint_IsLimitEnabled( b2JointId jointId )
{
	B2MI_JOINT_OF( jointId, b2_distanceJoint );
	return ( readJoint( w, joint ).flags & kJointLimit ) != 0;
}

B2_API void b2DistanceJoint_SetLengthRange( b2JointId jointId, float minLength, float maxLength )
{
	B2MI_JOINT_OF( jointId, b2_distanceJoint );
	minLength = b2ClampFloat( minLength, kLinearSlop, kHuge );
	maxLength = b2ClampFloat( maxLength, kLinearSlop, kHuge );
	DevDistanceJoint& j = editJoint( w, joint ).distance;
	j.minLength = b2MinFloat( minLength, maxLength );
	j.maxLength = b2MaxFloat( minLength, maxLength );
	j.impulse = 0.0f;
	j.lowerImpulse = 0.0f;
	j.upperImpulse = 0.0f;
}

B2_API float b2DistanceJoint_GetMinLength( b2JointId jointId )
{
	B2MI_JOINT_OF( jointId, b2_distanceJoint );
	return readJoint( w, joint ).distance.minLength;
}

B2_API float b2DistanceJoint_GetMaxLength( b2JointId jointId )
{
	B2MI_JOINT_OF( jointId, b2_distanceJoint );
	return readJoint( w, joint ).distance.maxLength;
}

B2_API float b2DistanceJoint_GetCurrentLength( b2JointId jointId )
{
	B2MI_JOINT_OF( jointId, b2_distanceJoint );
	const DevJoint& d = readJoint( w, joint );
	b2Transform xfA = b2Body_GetTransform( bodyIdOf( w, joint->edges[0].bodyId ) );
	b2Transform xfB = b2Body_GetTransform( bodyIdOf( w, joint->edges[1].bodyId ) );
	b2Vec2 pA = b2TransformPoint( xfA, d.localOriginAnchorA );
	b2Vec2 pB = b2TransformPoint( xfB, d.localOriginAnchorB );
	return b2Length( b2Sub( pB, pA ) );
}

B2_API void b2DistanceJoint_EnableMotor( b2JointId jointId, bool enableMotor )
{
	B2MI_JOINT_OF( jointId, b2_distanceJoint );
	if ( enableMotor != ( ( readJoint( w, joint ).flags & kJointMotor ) != 0 ) )
	{
		DevJoint& d = editJoint( w, joint );
		setFlag( d, kJointMotor, enableMotor );
		d.distance.motorImpulse = 0.0f;
	}
}

B2_API bool b2DistanceJoint_IsMotorEnabled( b2JointId jointId )
{
	B2MI_JOINT_OF( jointId, b2_distanceJoint );
	return ( readJoint( w, joint ).flags & kJointMotor ) != 0;
}

B2_API void b2DistanceJoint_SetMotorSpeed( b2JointId jointId, float motorSpeed )
{
	B2MI_JOINT_OF( jointId, b2_distanceJoint );
	editJoint( w, joint ).distance.motorSpeed = motorSpeed;
}

B2_API float b2DistanceJoint_GetMotorSpeed( b2JointId jointId )
{
	B2MI_JOINT_OF( jointId, b2_distanceJoint );
	return readJoint( w, joint ).distance.motorSpeed;
}

B2_API void b2DistanceJoint_SetMaxMotorForce( b2JointId jointId, float force )
{
	B2MI_JOINT_OF( jointId, b2_distanceJoint );
	editJoint( w, joint ).distance.maxMotorForce = force;
}

B2_API float b2DistanceJoint_GetMaxMotorForce( b2JointId jointId )
{
	B2MI_JOINT_OF( jointId, b2_distanceJoint );
	return readJoint( w, joint ).distance.maxMotorForce;
}

B2_API float b2DistanceJoint_GetMotorForce( b2JointId jointId )
{
	B2MI_JOINT_OF( jointId, b2_distanceJoint );
	return lastInvH( w ) * readJoint( w, joint ).distance.motorImpulse;
}

// ---- motor joint (src/motor_joint.c:14-72) ----

B2_API void b2MotorJoint_SetLinearOffset( b2JointId jointId, b2Vec2 linearOffset )
{
	B2MI_JOINT_OF( jointId, b2_motorJoint );
	editJoint( w, joint ).motor.linearOffset = linearOffset;
}

B2_API b2Vec2 b2MotorJoint_GetLinearOffset( b2JointId jointId )
{
	B2MI_JOINT_OF( jointId, b2_motorJoint );
	return readJoint( w, joint ).motor.linearOffset;
}

B2_API void b2MotorJoint_SetAngularOffset( b2JointId jointId, float angularOffset )
{
	B2MI_JOINT_OF( jointId, b2_motorJoint );
	editJoint( w, joint ).motor.angularOffset = b2ClampFloat( angularOffset, -b2_pi, b2_pi );
}

B2_API float b2MotorJoint_GetAngularOffset( b2JointId jointId )
{
	B2MI_JOINT_OF( jointId, b2_motorJoint );
	return readJoint( w, joint ).motor.angularOffset;
}

B2_API void b2MotorJoint_SetMaxForce( b2JointId jointId, float maxForce )
{
	B2MI_JOINT_OF( jointId, b2_motorJoint );
	editJoint( w, joint ).motor.maxForce = b2MaxFloat( 0.0f, maxForce );
}

B2_API float b2MotorJoint_GetMaxForce( b2JointId jointId )
{
	B2MI_JOINT_OF( jointId, b2_motorJoint );
	return readJoint( w, joint ).motor.maxForce;
}

B2_API void b2MotorJoint_SetMaxTorque( b2JointId jointId, float maxTorque )
{
	B2MI_JOINT_OF( jointId, b2_motorJoint );
	editJoint( w, joint ).motor.maxTorque = b2MaxFloat( 0.0f, maxTorque );
}

B2_API float b2MotorJoint_GetMaxTorque( b2JointId jointId )
{
	B2MI_JOINT_OF( jointId, b2_motorJoint );
	return readJoint( w, joint ).motor.maxTorque;
}

B2_API void b2MotorJoint_SetCorrectionFactor( b2JointId jointId, float correctionFactor )
{
	B2MI_JOINT_OF( jointId, b2_motorJoint );
	editJoint( w, joint ).motor.correctionFactor = b2ClampFloat( correctionFactor, 0.0f, 1.0f );
}

B2_API float b2MotorJoint_GetCorrectionFactor( b2JointId jointId )
{
	B2MI_JOINT_OF( jointId, b2_motorJoint );
	return readJoint( w, joint ).motor.correctionFactor;
}

// ---- mouse joint (src/mouse_joint.c:14-64) ----

B2_API void b2MouseJoint_SetTarget( b2JointId jointId, b2Vec2 target )
{
	B2MI_JOINT_OF( jointId, b2_mouseJoint );
	editJoint( w, joint ).mouse.targetA = target;
}

B2_API b2Vec2 b2MouseJoint_GetTarget( b2JointId jointId )
{
	B2MI_JOINT_OF( jointId, b2_mouseJoint );
	return readJoint( w, joint ).mouse.targetA;
}

B2_API void b2MouseJoint_SetSpringHertz( b2JointId jointId, float hertz )
{
	B2MI_JOINT_OF( jointId, b2_mouseJoint );
	editJoint( w, joint ).mouse.hertz = hertz;
}

B2_API float b2MouseJoint_GetSpringHertz( b2JointId jointId )
{
	B2MI_JOINT_OF( jointId, b2_mouseJoint );
	return readJoint( w, joint ).mouse.hertz;
}

B2_API void b2MouseJoint_SetSpringDampingRatio( b2JointId jointId, float dampingRatio )
{
	B2MI_JOINT_OF( jointId, b2_mouseJoint );
	editJoint( w, joint ).mouse.dampingRatio = dampingRatio;
}

B2_API float b2MouseJoint_GetSpringDampingRatio( b2JointId jointId )
{
	B2MI_JOINT_OF( jointId, b2_mouseJoint );
	return readJoint( w, joint ).mouse.dampingRatio;
}

B2_API void b2MouseJoint_SetMaxForce( b2JointId jointId, float maxForce )
{
	B2MI_JOINT_OF( jointId, b2_mouseJoint );
	editJoint( w, joint ).mouse.maxForce = maxForce;
}

B2_API float b2MouseJoint_GetMaxForce( b2JointId jointId )
{
	B2MI_JOINT_OF( jointId, b2_mouseJoint );
	return readJoint( w, joint ).mouse.maxForce;
}

// ---- prismatic joint (src/prismatic_joint.c:16-142) ----

B2_API void b2PrismaticJoint_EnableSpring( b2JointId jointId, bool enableSpring )
{
	B2MI_JOINT_OF( jointId, b2_prismaticJoint );
	if ( enableSpring != ( ( readJoint( w, joint ).flags & kJointSpring ) != 0 ) )
	{
		DevJoint& d = editJoint( w, joint );
		setFlag( d, kJointSpring, enableSpring );
		d.prismatic.springImpulse = 0.0f;
	}
}

B2_API bool b2PrismaticJoint_IsSpringEnabled( b2JointId jointId )
{
	B2MI_JOINT_OF( jointId, b2_prismaticJoint );
	return ( readJoint( w, joint ).flags & kJointSpring ) != 0;
}

B2_API void b2PrismaticJoint_SetSpringHertz( b2JointId jointId, float hertz )
{
	B2MI_JOINT_OF( jointId, b2_prismaticJoint );
	editJoint( w, joint ).prismatic.hertz = hertz;
}

B2_API float b2PrismaticJoint_GetSpringHertz( b2JointId jointId )
{
	B2MI_JOINT_OF( jointId, b2_prismaticJoint );
	return readJoint( w, joint ).prismatic.hertz;
}

B2_API void b2PrismaticJoint_SetSpringDampingRatio( b2JointId jointId, float dampingRatio )
{
	B2MI_JOINT_OF( jointId, b2_prismaticJoint );
	editJoint( w, joint ).prismatic.dampingRatio = dampingRatio;
}

B2_API float b2PrismaticJoint_GetSpringDampingRatio( b2JointId jointId )
{
	B2MI_JOINT_OF( jointId, b2_prismaticJoint );
	return readJoint( w, joint ).prismatic.dampingRatio;
}

B2_API void b2PrismaticJoint_EnableLimit( b2JointId jointId, bool enableLimit )
{
	B2MI_JOINT_OF( jointId, b2_prismaticJoint );
	if ( enableLimit != ( ( readJoint( w, joint ).flags & kJointLimit ) != 0 ) )
	{
		DevJoint& d = editJoint( w, joint );
		setFlag( d, kJointLimit, enableLimit );
		d.prismatic.lowerImpulse = 0.0f;
		d.prismatic.upperImpulse = 0.0f;
	}
}

B2_API bool b2PrismaticJoint_IsLimitEnabled( b2JointId jointId )
{
	B2MI_JOINT_OF( jointId, b2_prismaticJoint );
	return ( readJoint( w, joint ).flags & kJointLimit ) != 0;
}

B2_API float b2PrismaticJoint_GetLowerLimit( b2JointId jointId )
{
	B2MI_JOINT_OF( jointId, b2_prismaticJoint );
	return readJoint( w, joint ).prismatic.lowerTranslation;
}

B2_API float b2PrismaticJoint_GetUpperLimit( b2JointId jointId )
{
	B2MI_JOINT_OF( jointId, b2_prismaticJoint );
	return readJoint( w, joint ).prismatic.upperTranslation;
}

B2_API void b2PrismaticJoint_SetLimits( b2JointId jointId, float lower, float upper )
{
	B2MI_JOINT_OF( jointId, b2_prismaticJoint );
	const DevPrismaticJoint& r = readJoint( w, joint ).prismatic;
	if ( lower != r.lowerTranslation || upper != r.upperTranslation )
	{
		DevPrismaticJoint& e = editJoint( w, joint ).prismatic;
		e.lowerTranslation = b2MinFloat( lower, upper );
		e.upperTranslation = b2MaxFloat( lower, upper );
		e.lowerImpulse = 0.0f;
		e.upperImpulse = 0.0f;
	}
}

B2_API void b2PrismaticJoint_EnableMotor( b2JointId jointId, bool enableMotor )
{
	B2MI_JOINT_OF( jointId, b2_prismaticJoint );
	if ( enableMotor != ( ( readJoint( w, joint ).flags & kJointMotor ) != 0 ) )
	{
		DevJoint& d = editJoint( w, joint );
		setFlag( d, kJointMotor, enableMotor );
		d.prismatic.motorImpulse = 0.0f;
	}
}

B2_API bool b2PrismaticJoint_IsMotorEnabled( b2JointId jointId )
{
	B2MI_JOINT_OF( jointId, b2_prismaticJoint );
	return ( readJoint( w, joint ).flags & kJointMotor ) != 0;
}

B2_API void b2PrismaticJoint_SetMotorSpeed( b2JointId jointId, float motorSpeed )
{
	B2MI_JOINT_OF( jointId, b2_prismaticJoint );
	editJoint( w, joint ).prismatic.motorSpeed = motorSpeed;
}

B2_API float b2PrismaticJoint_GetMotorSpeed( b2JointId jointId )
{
	B2MI_JOINT_OF( jointId, b2_prismaticJoint );
	return readJoint( w, joint ).prismatic.motorSpeed;
}

B2_API void b2PrismaticJoint_SetMaxMotorForce( b2JointId jointId, float force )
{
	B2MI_JOINT_OF( jointId, b2_prismaticJoint );
	editJoint( w, joint ).prismatic.maxMotorForce = force;
}

B2_API float b2PrismaticJoint_GetMaxMotorForce( b2JointId jointId )
{
	B2MI_JOINT_OF( jointId, b2_prismaticJoint );
	return readJoint( w, joint ).prismatic.maxMotorForce;
}

B2_API float b2PrismaticJoint_GetMotorForce( b2JointId jointId )
{
	B2MI_JOINT_OF( jointId, b2_prismaticJoint );
	return lastInvH( w ) * readJoint( w, joint ).prismatic.motorImpulse;
}

// ---- wheel joint (src/wheel_joint.c:16-143) ----

B2_API void b2WheelJoint_EnableSpring( b2JointId jointId, bool enableSpring )
{
	B2MI_JOINT_OF( jointId, b2_wheelJoint );
	if ( enableSpring != ( ( readJoint( w, joint ).flags & kJointSpring ) != 0 ) )
	{
		DevJoint& d = editJoint( w, joint );
		setFlag( d, kJointSpring, enableSpring );
		d.wheel.springImpulse = 0.0f;
	}
}

B2_API bool b2WheelJoint_IsSpringEnabled( b2JointId jointId )
{
	B2MI_JOINT_OF( jointId, b2_wheelJoint );
	return ( readJoint( w, joint ).flags & kJointSpring ) != 0;
}

B2_API void b2WheelJoint_SetSpringHertz( b2JointId jointId, float hertz )
{
	B2MI_JOINT_OF( jointId, b2_wheelJoint );
	editJoint( w, joint ).wheel.hertz = hertz;
}

B2_API float b2WheelJoint_GetSpringHertz( b2JointId jointId )
{
	B2MI_JOINT_OF( jointId, b2_wheelJoint );
	return readJoint( w, joint ).wheel.hertz;
}

B2_API void b2WheelJoint_SetSpringDampingRatio( b2JointId jointId, float dampingRatio )
{
	B2MI_JOINT_OF( jointId, b2_wheelJoint );
	editJoint( w, joint ).wheel.dampingRatio = dampingRatio;
}

B2_API float b2WheelJoint_GetSpringDampingRatio( b2JointId jointId )
{
	B2MI_JOINT_OF( jointId, b2_wheelJoint );
	return readJoint( w, joint ).wheel.dampingRatio;
}

B2_API void b2WheelJoint_EnableLimit( b2JointId jointId, bool enableLimit )
{
	B2MI_JOINT_OF( jointId, b2_wheelJoint );
	if ( enableLimit != ( ( readJoint( w, joint ).flags & kJointLimit ) != 0 ) )
	{
		DevJoint& d = editJoint( w, joint );
		setFlag( d, kJointLimit, enableLimit );
		d.wheel.lowerImpulse = 0.0f;
		d.wheel.upperImpulse = 0.0f;
	}
}

B2_API bool b2WheelJoint_IsLimitEnabled( b2JointId jointId )
{
	B2MI_JOINT_OF( jointId, b2_wheelJoint );
	return ( readJoint( w, joint ).flags & kJointLimit ) != 0;
}

B2_API float b2WheelJoint_GetLowerLimit( b2JointId jointId )
{
	B2MI_JOINT_OF( jointId, b2_wheelJoint );
	return readJoint( w, joint ).wheel.lowerTranslation;
}

B2_API float b2WheelJoint_GetUpperLimit( b2JointId jointId )
{
	B2MI_JOINT_OF( jointId, b2_wheelJoint );
	return readJoint( w, joint ).wheel.upperTranslation;
}

B2_API void b2WheelJoint_SetLimits( b2JointId jointId, float lower, float upper )
{
	B2MI_JOINT_OF( jointId, b2_wheelJoint );
	const DevWheelJoint& r = readJoint( w, joint ).wheel;
	if ( lower != r.lowerTranslation || upper != r.upperTranslation )
	{
		DevWheelJoint& e = editJoint( w, joint ).wheel;
		e.lowerTranslation = b2MinFloat( lower, upper );
		e.upperTranslation = b2MaxFloat( lower, upper );
		e.lowerImpulse = 0.0f;
		e.upperImpulse = 0.0f;
	}
}

B2_API void b2WheelJoint_EnableMotor( b2JointId jointId, bool enableMotor )
{
	B2MI_JOINT_OF( jointId, b2_wheelJoint );
	if ( enableMotor != ( ( readJoint( w, joint ).flags & kJointMotor ) != 0 ) )
	{
		DevJoint& d = editJoint( w, joint );
		setFlag( d, kJointMotor, enableMotor );
		d.wheel.motorImpulse = 0.0f;
	}
}

B2_API bool b2WheelJoint_IsMotorEnabled( b2JointId jointId )
{
	B2MI_JOINT_OF( jointId, b2_wheelJoint );
	return ( readJoint( w, joint ).flags & kJointMotor ) != 0;
}

B2_API void b2WheelJoint_SetMotorSpeed( b2JointId jointId, float motorSpeed )
{
	B2MI_JOINT_OF( jointId, b2_wheelJoint );
	editJoint( w, joint ).wheel.motorSpeed = motorSpeed;
}

B2_API float b2WheelJoint_GetMotorSpeed( b2JointId jointId )
{
	B2MI_JOINT_OF( jointId, b2_wheelJoint );
	return readJoint( w, joint ).wheel.motorSpeed;
}

B2_API void b2WheelJoint_SetMaxMotorTorque( b2JointId jointId, float torque )
{
	B2MI_JOINT_OF( jointId, b2_wheelJoint );
	editJoint( w, joint ).wheel.maxMotorTorque = torque;
}

B2_API float b2WheelJoint_GetMaxMotorTorque( b2JointId jointId )
{
	B2MI_JOINT_OF( jointId, b2_wheelJoint );
	return readJoint( w, joint ).wheel.maxMotorTorque;
}

B2_API float b2WheelJoint_GetMotorTorque( b2JointId jointId )
{
	B2MI_JOINT_OF( jointId, b2_wheelJoint );
	return lastInvH( w ) * readJoint( w, joint ).wheel.motorImpulse;
}

} // extern "C"
