// CPU ORACLE — test infrastructure only (see or_internal.h).
// World / body / shape / contact model, broadphase pair update, graph colouring and the
// narrowphase driver, restating src/world.c, src/body.c, src/shape.c, src/contact.c,
// src/broad_phase.c and src/constraint_graph.c (line references at each function).
#include "or_internal.h"

#include <float.h>
#include <math.h>

orWorld or_worlds[OR_MAX_WORLDS];

// ---- defaults (src/types.c:8-60) ----
B2_API b2WorldDef b2DefaultWorldDef( void )
{
	b2WorldDef def = { 0 };
	def.gravity = ( b2Vec2 ){ 0.0f, -10.0f };
	def.hitEventThreshold = 1.0f;
	def.restitutionThreshold = 1.0f;
	def.contactPushoutVelocity = 3.0f;
	def.contactHertz = 30.0f;
	def.contactDampingRatio = 10.0f;
	def.jointHertz = 60.0f;
	def.jointDampingRatio = 2.0f;
	def.enableSleep = true;
	def.enableContinous = true;
	def.internalValue = OR_COOKIE;
	return def;
}

B2_API b2BodyDef b2DefaultBodyDef( void )
{
	b2BodyDef def = { 0 };
	def.type = b2_staticBody;
	def.rotation = b2Rot_identity;
	def.sleepThreshold = 0.05f;
	def.gravityScale = 1.0f;
	def.enableSleep = true;
	def.isAwake = true;
	def.isEnabled = true;
	def.automaticMass = true;
	def.internalValue = OR_COOKIE;
	return def;
}

B2_API b2Filter b2DefaultFilter( void )
{
	b2Filter f = { 0x00000001, 0xFFFFFFFF, 0 };
	return f;
}

B2_API b2ShapeDef b2DefaultShapeDef( void )
{
	b2ShapeDef def = { 0 };
	def.friction = 0.6f;
	def.density = 1.0f;
	def.filter = b2DefaultFilter();
	def.enableSensorEvents = true;
	def.enableContactEvents = true;
	def.internalValue = OR_COOKIE;
	return def;
}

// ---- lookups ----
orWorld* or_get_world( b2WorldId id )
{
	OR_ASSERT( 1 <= id.index1 && id.index1 <= OR_MAX_WORLDS );
	orWorld* world = or_worlds + ( id.index1 - 1 );
	OR_ASSERT( world->worldId == id.index1 - 1 && world->revision == id.revision );
	return world;
}

orWorld* or_get_world_index( int index )
{
	OR_ASSERT( 0 <= index && index < OR_MAX_WORLDS );
	orWorld* world = or_worlds + index;
	OR_ASSERT( world->worldId == index );
	return world;
}

orBody* or_get_body_full( orWorld* world, b2BodyId id )
{
	int index = id.index1 - 1;
	OR_ASSERT( 0 <= index && index < world->bodies.count );
	orBody* body = world->bodies.data + index;
	OR_ASSERT( body->id == index && body->revision == id.revision );
	return body;
}

orBodySim* or_get_body_sim( orWorld* world, orBody* body )
{
	if ( body->setIndex == 0 )
		return world->staticSims.data + body->localIndex;
	return world->awakeSims.data + body->localIndex;
}

static orBodyState* or_get_body_state( orWorld* world, orBody* body )
{
	if ( body->setIndex == 2 )
		return world->awakeStates.data + body->localIndex;
	return NULL;
}

orShape* or_get_shape_full( orWorld* world, b2ShapeId id )
{
	int index = id.index1 - 1;
	OR_ASSERT( 0 <= index && index < world->shapes.count );
	orShape* shape = world->shapes.data + index;
	OR_ASSERT( shape->id == index && shape->revision == id.revision );
	return shape;
}

// ---- world (src/world.c:80-288) ----
B2_API b2WorldId b2CreateWorld( const b2WorldDef* def )
{
	OR_ASSERT( def->internalValue == OR_COOKIE );
	int worldId = OR_NULL;
	for ( int i = 0; i < OR_MAX_WORLDS; ++i )
	{
		if ( or_worlds[i].inUse == false )
		{
			worldId = i;
			break;
		}
	}
	if ( worldId == OR_NULL )
		return ( b2WorldId ){ 0 };

	orWorld* world = or_worlds + worldId;
	uint16_t revision = world->revision;
	memset( world, 0, sizeof( *world ) );
	world->worldId = worldId;
	world->revision = revision;
	world->inUse = true;
	for ( int i = 0; i < 3; ++i )
		or_tree_create( world->trees + i );
	for ( int i = 0; i < OR_GRAPH_COLORS - 1; ++i )
		or_bitset_set_count_and_clear( &world->colors[i].bodySet, 16 );

	world->gravity = def->gravity;
	world->hitEventThreshold = def->hitEventThreshold;
	world->restitutionThreshold = def->restitutionThreshold;
	world->contactPushoutVelocity = def->contactPushoutVelocity;
	world->contactHertz = def->contactHertz;
	world->contactDampingRatio = def->contactDampingRatio;
	world->jointHertz = def->jointHertz;
	world->jointDampingRatio = def->jointDampingRatio;
	world->enableSleep = def->enableSleep;
	world->enableWarmStarting = true;
	world->enableContinuous = def->enableContinous;
	return ( b2WorldId ){ (uint16_t)( worldId + 1 ), world->revision };
}

B2_API void b2DestroyWorld( b2WorldId worldId )
{
	orWorld* world = or_get_world( worldId );
	OR_FREE( world->bodies );
	OR_FREE( world->bodyIdPool.freeArray );
	OR_FREE( world->staticSims );
	OR_FREE( world->awakeSims );
	OR_FREE( world->awakeStates );
	OR_FREE( world->awakeContacts );
	OR_FREE( world->shapes );
	OR_FREE( world->shapeIdPool.freeArray );
	OR_FREE( world->contacts );
	OR_FREE( world->contactIdPool.freeArray );
	OR_FREE( world->joints );
	OR_FREE( world->jointIdPool.freeArray );
	OR_FREE( world->staticJoints );
	for ( int i = 0; i < OR_GRAPH_COLORS; ++i )
	{
		or_bitset_free( &world->colors[i].bodySet );
		OR_FREE( world->colors[i].contacts );
		OR_FREE( world->colors[i].joints );
	}
	for ( int i = 0; i < 3; ++i )
		or_tree_destroy( world->trees + i );
	or_bitset_free( &world->moveSet );
	OR_FREE( world->moveArray );
	or_keyset_free( &world->pairSet );
	OR_FREE( world->moveEvents );
	OR_FREE( world->sensorBegin );
	OR_FREE( world->sensorEnd );
	OR_FREE( world->contactBegin );
	OR_FREE( world->contactEnd );
	OR_FREE( world->contactHit );
	uint16_t revision = world->revision;
	memset( world, 0, sizeof( *world ) );
	world->worldId = OR_NULL;
	world->revision = revision + 1;
}

B2_API bool b2World_IsValid( b2WorldId id )
{
	if ( id.index1 < 1 || OR_MAX_WORLDS < id.index1 )
		return false;
	orWorld* world = or_worlds + ( id.index1 - 1 );
	if ( world->inUse == false || world->worldId != id.index1 - 1 )
		return false;
	return id.revision == world->revision;
}

B2_API void b2World_EnableSleeping( b2WorldId worldId, bool flag )
{
	or_get_world( worldId )->enableSleep = flag;
}
B2_API void b2World_EnableContinuous( b2WorldId worldId, bool flag )
{
	or_get_world( worldId )->enableContinuous = flag;
}
B2_API void b2World_SetRestitutionThreshold( b2WorldId worldId, float value )
{
	or_get_world( worldId )->restitutionThreshold = b2ClampFloat( value, 0.0f, FLT_MAX );
}
B2_API void b2World_SetHitEventThreshold( b2WorldId worldId, float value )
{
	or_get_world( worldId )->hitEventThreshold = b2ClampFloat( value, 0.0f, FLT_MAX );
}
B2_API void b2World_SetCustomFilterCallback( b2WorldId worldId, b2CustomFilterFcn* fcn, void* context )
{
	(void)worldId;
	(void)context;
	OR_ASSERT( fcn == NULL && "custom filter callbacks are not supported" );
}
B2_API void b2World_SetPreSolveCallback( b2WorldId worldId, b2PreSolveFcn* fcn, void* context )
{
	(void)worldId;
	(void)context;
	OR_ASSERT( fcn == NULL && "pre-solve callbacks are not supported" );
}
B2_API void b2World_SetGravity( b2WorldId worldId, b2Vec2 gravity )
{
	or_get_world( worldId )->gravity = gravity;
}
B2_API b2Vec2 b2World_GetGravity( b2WorldId worldId )
{
	return or_get_world( worldId )->gravity;
}
// src/world.c:1629-1636
B2_API void b2World_SetContactTuning( b2WorldId worldId, float hertz, float dampingRatio, float pushVelocity )
{
	orWorld* world = or_get_world( worldId );
	world->contactHertz = b2ClampFloat( hertz, 0.0f, FLT_MAX );
	world->contactDampingRatio = b2ClampFloat( dampingRatio, 0.0f, FLT_MAX );
	world->contactPushoutVelocity = b2ClampFloat( pushVelocity, 0.0f, FLT_MAX );
}
B2_API void b2World_EnableWarmStarting( b2WorldId worldId, bool flag )
{
	or_get_world( worldId )->enableWarmStarting = flag;
}
B2_API b2Profile b2World_GetProfile( b2WorldId worldId )
{
	return or_get_world( worldId )->profile;
}

static int or_id_count( const orIdPool* pool )
{
	return pool->nextIndex - pool->freeArray.count;
}

// src/world.c:1666-1692
B2_API b2Counters b2World_GetCounters( b2WorldId worldId )
{
	orWorld* world = or_get_world( worldId );
	b2Counters s = { 0 };
	s.staticBodyCount = world->staticSims.count;
	s.bodyCount = or_id_count( &world->bodyIdPool );
	s.shapeCount = or_id_count( &world->shapeIdPool );
	s.contactCount = or_id_count( &world->contactIdPool );
	s.jointCount = or_id_count( &world->jointIdPool );
	s.staticTreeHeight = or_tree_height( world->trees + b2_staticBody );
	s.treeHeight = b2MaxInt( or_tree_height( world->trees + b2_dynamicBody ), or_tree_height( world->trees + b2_kinematicBody ) );
	for ( int i = 0; i < OR_GRAPH_COLORS; ++i )
		s.colorCounts[i] = world->colors[i].contacts.count + world->colors[i].joints.count;
	return s;
}

// ---- bodies (src/body.c:178-318, 516-617) ----
B2_API b2BodyId b2CreateBody( b2WorldId worldId, const b2BodyDef* def )
{
	OR_ASSERT( def->internalValue == OR_COOKIE );
	orWorld* world = or_get_world( worldId );
	if ( world->locked )
		return b2_nullBodyId;
	OR_ASSERT( def->isEnabled && "disabled bodies are not supported" );
	OR_ASSERT( ( def->isAwake || def->enableSleep == false || def->type == b2_staticBody ) &&
			   "bodies created asleep are not supported" );

	int setId = def->type == b2_staticBody ? 0 : 2;
	int bodyId = or_alloc_id( &world->bodyIdPool );

	orBodySim sim = { 0 };
	sim.transform.p = def->position;
	sim.transform.q = def->rotation;
	sim.center = def->position;
	sim.rotation0 = sim.transform.q;
	sim.center0 = sim.center;
	sim.minExtent = OR_HUGE;
	sim.linearDamping = def->linearDamping;
	sim.angularDamping = def->angularDamping;
	sim.gravityScale = def->gravityScale;
	sim.bodyId = bodyId;
	sim.isBullet = def->isBullet;
#if ORACLE_SOLVER_F64
	sim.centerR = ( orVec2r ){ def->position.x, def->position.y };
	sim.qR = ( orVec2r ){ def->rotation.c, def->rotation.s };
#endif
	int localIndex;
	if ( setId == 0 )
	{
		localIndex = world->staticSims.count;
		OR_PUSH( world->staticSims, sim );
	}
	else
	{
		localIndex = world->awakeSims.count;
		OR_PUSH( world->awakeSims, sim );
		orBodyState st = { 0 };
		st.v = ( orVec2r ){ def->linearVelocity.x, def->linearVelocity.y };
		st.w = def->angularVelocity;
		st.dq = ( orVec2r ){ 1, 0 };
		OR_PUSH( world->awakeStates, st );
	}

	if ( bodyId == world->bodies.count )
	{
		orBody empty = { 0 };
		OR_PUSH( world->bodies, empty );
	}
	orBody* body = world->bodies.data + bodyId;
	uint16_t revision = body->revision;
	memset( body, 0, sizeof( *body ) );
	body->userData = def->userData;
	body->setIndex = setId;
	body->localIndex = localIndex;
	body->revision = revision + 1;
	body->headShapeId = OR_NULL;
	body->headContactKey = OR_NULL;
	body->headJointKey = OR_NULL;
	body->jointCount = 0;
	body->id = bodyId;
	body->sleepThreshold = def->sleepThreshold;
	body->type = def->type;
	body->enableSleep = def->enableSleep;
	body->fixedRotation = def->fixedRotation;
	body->automaticMass = def->automaticMass;
	return ( b2BodyId ){ bodyId + 1, (uint16_t)world->worldId, body->revision };
}

static b2MassData or_compute_shape_mass( const orShape* s )
{
	switch ( s->type )
	{
		case b2_capsuleShape:
			return b2ComputeCapsuleMass( &s->capsule, s->density );
		case b2_circleShape:
			return b2ComputeCircleMass( &s->circle, s->density );
		case b2_polygonShape:
			return b2ComputePolygonMass( &s->polygon, s->density );
		default:
			return ( b2MassData ){ 0 };
	}
}

// src/shape.c:551-619
static void or_compute_shape_extent( const orShape* s, b2Vec2 localCenter, float* minExtent, float* maxExtent )
{
	switch ( s->type )
	{
		case b2_capsuleShape:
		{
			float radius = s->capsule.radius;
			*minExtent = radius;
			b2Vec2 c1 = b2Sub( s->capsule.center1, localCenter );
			b2Vec2 c2 = b2Sub( s->capsule.center2, localCenter );
			*maxExtent = sqrtf( b2MaxFloat( b2LengthSquared( c1 ), b2LengthSquared( c2 ) ) ) + radius;
		}
		break;
		case b2_circleShape:
			*minExtent = s->circle.radius;
			*maxExtent = b2Length( b2Sub( s->circle.center, localCenter ) ) + s->circle.radius;
			break;
		case b2_polygonShape:
		{
			const b2Polygon* poly = &s->polygon;
			float minE = OR_HUGE, maxSqr = 0.0f;
			for ( int i = 0; i < poly->count; ++i )
			{
				b2Vec2 v = poly->vertices[i];
				float planeOffset = b2Dot( poly->normals[i], b2Sub( v, poly->centroid ) );
				minE = b2MinFloat( minE, planeOffset );
				maxSqr = b2MaxFloat( maxSqr, b2LengthSquared( b2Sub( v, localCenter ) ) );
			}
			*minExtent = minE + poly->radius;
			*maxExtent = sqrtf( maxSqr ) + poly->radius;
		}
		break;
		case b2_segmentShape:
		{
			*minExtent = 0.0f;
			b2Vec2 c1 = b2Sub( s->segment.point1, localCenter );
			b2Vec2 c2 = b2Sub( s->segment.point2, localCenter );
			*maxExtent = sqrtf( b2MaxFloat( b2LengthSquared( c1 ), b2LengthSquared( c2 ) ) );
		}
		break;
		default:
			*minExtent = 0.0f;
			*maxExtent = 0.0f;
			break;
	}
}

// src/body.c:516-617
void or_update_body_mass( orWorld* world, orBody* body )
{
	orBodySim* sim = or_get_body_sim( world, body );
	sim->mass = 0.0f;
	sim->invMass = 0.0f;
	sim->inertia = 0.0f;
	sim->invInertia = 0.0f;
	sim->localCenter = b2Vec2_zero;
	sim->minExtent = OR_HUGE;
	sim->maxExtent = 0.0f;

	if ( body->type != b2_dynamicBody )
	{
		sim->center = sim->transform.p;
#if ORACLE_SOLVER_F64
		sim->centerR = ( orVec2r ){ sim->center.x, sim->center.y };
#endif
		if ( body->type == b2_kinematicBody )
		{
			for ( int id = body->headShapeId; id != OR_NULL; id = world->shapes.data[id].nextShapeId )
			{
				float mn, mx;
				or_compute_shape_extent( world->shapes.data + id, b2Vec2_zero, &mn, &mx );
				sim->minExtent = b2MinFloat( sim->minExtent, mn );
				sim->maxExtent = b2MaxFloat( sim->maxExtent, mx );
			}
		}
		return;
	}

	b2Vec2 localCenter = b2Vec2_zero;
	for ( int id = body->headShapeId; id != OR_NULL; )
	{
		const orShape* s = world->shapes.data + id;
		id = s->nextShapeId;
		if ( s->density == 0.0f )
			continue;
		b2MassData md = or_compute_shape_mass( s );
		sim->mass += md.mass;
		localCenter = b2MulAdd( localCenter, md.mass, md.center );
		sim->inertia += md.rotationalInertia;
	}
	if ( sim->mass > 0.0f )
	{
		sim->invMass = 1.0f / sim->mass;
		localCenter = b2MulSV( sim->invMass, localCenter );
	}
	if ( sim->inertia > 0.0f && body->fixedRotation == false )
	{
		sim->inertia -= sim->mass * b2Dot( localCenter, localCenter );
		sim->invInertia = 1.0f / sim->inertia;
	}
	else
	{
		sim->inertia = 0.0f;
		sim->invInertia = 0.0f;
	}

	b2Vec2 oldCenter = sim->center;
	sim->localCenter = localCenter;
	sim->center = b2TransformPoint( sim->transform, sim->localCenter );
#if ORACLE_SOLVER_F64
	sim->centerR = ( orVec2r ){ sim->center.x, sim->center.y };
#endif
	orBodyState* state = or_get_body_state( world, body );
	if ( state != NULL )
	{
		// v += w x (c - c_old), in float like the reference
		float w = (float)state->w;
		b2Vec2 deltaLinear = b2CrossSV( w, b2Sub( sim->center, oldCenter ) );
		state->v.x = (float)state->v.x + deltaLinear.x;
		state->v.y = (float)state->v.y + deltaLinear.y;
	}

	for ( int id = body->headShapeId; id != OR_NULL; id = world->shapes.data[id].nextShapeId )
	{
		float mn, mx;
		or_compute_shape_extent( world->shapes.data + id, localCenter, &mn, &mx );
		sim->minExtent = b2MinFloat( sim->minExtent, mn );
		sim->maxExtent = b2MaxFloat( sim->maxExtent, mx );
	}
}

B2_API bool b2Body_IsValid( b2BodyId id )
{
	if ( id.world0 >= OR_MAX_WORLDS )
		return false;
	orWorld* world = or_worlds + id.world0;
	if ( world->inUse == false || world->worldId != id.world0 )
		return false;
	if ( id.index1 < 1 || world->bodies.count < id.index1 )
		return false;
	orBody* body = world->bodies.data + ( id.index1 - 1 );
	if ( body->setIndex == OR_NULL || body->id != id.index1 - 1 )
		return false;
	return body->revision == id.revision;
}

#define OR_BODY( bodyId ) \
	orWorld* world = or_get_world_index( ( bodyId ).world0 ); \
	orBody* body = or_get_body_full( world, ( bodyId ) )

B2_API b2BodyType b2Body_GetType( b2BodyId bodyId )
{
	OR_BODY( bodyId );
	return body->type;
}
B2_API void b2Body_SetUserData( b2BodyId bodyId, void* userData )
{
	OR_BODY( bodyId );
	body->userData = userData;
}
B2_API void* b2Body_GetUserData( b2BodyId bodyId )
{
	OR_BODY( bodyId );
	return body->userData;
}
B2_API b2Vec2 b2Body_GetPosition( b2BodyId bodyId )
{
	OR_BODY( bodyId );
	return or_get_body_sim( world, body )->transform.p;
}
B2_API b2Rot b2Body_GetRotation( b2BodyId bodyId )
{
	OR_BODY( bodyId );
	return or_get_body_sim( world, body )->transform.q;
}
B2_API b2Transform b2Body_GetTransform( b2BodyId bodyId )
{
	OR_BODY( bodyId );
	return or_get_body_sim( world, body )->transform;
}
B2_API b2Vec2 b2Body_GetLocalPoint( b2BodyId bodyId, b2Vec2 worldPoint )
{
	return b2InvTransformPoint( b2Body_GetTransform( bodyId ), worldPoint );
}
B2_API b2Vec2 b2Body_GetWorldPoint( b2BodyId bodyId, b2Vec2 localPoint )
{
	return b2TransformPoint( b2Body_GetTransform( bodyId ), localPoint );
}
B2_API b2Vec2 b2Body_GetLinearVelocity( b2BodyId bodyId )
{
	OR_BODY( bodyId );
	orBodyState* st = or_get_body_state( world, body );
	return st ? ( b2Vec2 ){ (float)st->v.x, (float)st->v.y } : b2Vec2_zero;
}
B2_API float b2Body_GetAngularVelocity( b2BodyId bodyId )
{
	OR_BODY( bodyId );
	orBodyState* st = or_get_body_state( world, body );
	return st ? (float)st->w : 0.0f;
}
B2_API void b2Body_SetLinearVelocity( b2BodyId bodyId, b2Vec2 v )
{
	OR_BODY( bodyId );
	orBodyState* st = or_get_body_state( world, body );
	if ( st )
		st->v = ( orVec2r ){ v.x, v.y };
}
B2_API void b2Body_SetAngularVelocity( b2BodyId bodyId, float w )
{
	OR_BODY( bodyId );
	orBodyState* st = or_get_body_state( world, body );
	if ( st )
		st->w = w;
}
// src/body.c:799-904
B2_API void b2Body_ApplyForce( b2BodyId bodyId, b2Vec2 force, b2Vec2 point, bool wake )
{
	(void)wake;
	OR_BODY( bodyId );
	if ( body->setIndex != 2 )
		return;
	orBodySim* sim = or_get_body_sim( world, body );
	sim->force = b2Add( sim->force, force );
	sim->torque += b2Cross( b2Sub( point, sim->center ), force );
}
B2_API void b2Body_ApplyForceToCenter( b2BodyId bodyId, b2Vec2 force, bool wake )
{
	(void)wake;
	OR_BODY( bodyId );
	if ( body->setIndex != 2 )
		return;
	orBodySim* sim = or_get_body_sim( world, body );
	sim->force = b2Add( sim->force, force );
}
B2_API void b2Body_ApplyTorque( b2BodyId bodyId, float torque, bool wake )
{
	(void)wake;
	OR_BODY( bodyId );
	if ( body->setIndex != 2 )
		return;
	or_get_body_sim( world, body )->torque += torque;
}
B2_API void b2Body_ApplyLinearImpulse( b2BodyId bodyId, b2Vec2 impulse, b2Vec2 point, bool wake )
{
	(void)wake;
	OR_BODY( bodyId );
	if ( body->setIndex != 2 )
		return;
	orBodySim* sim = or_get_body_sim( world, body );
	orBodyState* st = or_get_body_state( world, body );
	b2Vec2 v = { (float)st->v.x, (float)st->v.y };
	v = b2MulAdd( v, sim->invMass, impulse );
	st->v = ( orVec2r ){ v.x, v.y };
	st->w = (float)st->w + sim->invInertia * b2Cross( b2Sub( point, sim->center ), impulse );
}
B2_API void b2Body_ApplyLinearImpulseToCenter( b2BodyId bodyId, b2Vec2 impulse, bool wake )
{
	(void)wake;
	OR_BODY( bodyId );
	if ( body->setIndex != 2 )
		return;
	orBodySim* sim = or_get_body_sim( world, body );
	orBodyState* st = or_get_body_state( world, body );
	b2Vec2 v = { (float)st->v.x, (float)st->v.y };
	v = b2MulAdd( v, sim->invMass, impulse );
	st->v = ( orVec2r ){ v.x, v.y };
}
B2_API void b2Body_ApplyAngularImpulse( b2BodyId bodyId, float impulse, bool wake )
{
	(void)wake;
	OR_BODY( bodyId );
	if ( body->setIndex != 2 )
		return;
	orBodySim* sim = or_get_body_sim( world, body );
	orBodyState* st = or_get_body_state( world, body );
	st->w = (float)st->w + sim->invInertia * impulse;
}
B2_API float b2Body_GetMass( b2BodyId bodyId )
{
	OR_BODY( bodyId );
	return or_get_body_sim( world, body )->mass;
}
B2_API float b2Body_GetInertiaTensor( b2BodyId bodyId )
{
	OR_BODY( bodyId );
	return or_get_body_sim( world, body )->inertia;
}
B2_API b2Vec2 b2Body_GetLocalCenterOfMass( b2BodyId bodyId )
{
	OR_BODY( bodyId );
	return or_get_body_sim( world, body )->localCenter;
}
B2_API b2Vec2 b2Body_GetWorldCenterOfMass( b2BodyId bodyId )
{
	OR_BODY( bodyId );
	return or_get_body_sim( world, body )->center;
}
B2_API b2MassData b2Body_GetMassData( b2BodyId bodyId )
{
	OR_BODY( bodyId );
	orBodySim* sim = or_get_body_sim( world, body );
	return ( b2MassData ){ sim->mass, sim->localCenter, sim->inertia };
}
B2_API void b2Body_SetLinearDamping( b2BodyId bodyId, float v )
{
	OR_BODY( bodyId );
	or_get_body_sim( world, body )->linearDamping = v;
}
B2_API float b2Body_GetLinearDamping( b2BodyId bodyId )
{
	OR_BODY( bodyId );
	return or_get_body_sim( world, body )->linearDamping;
}
B2_API void b2Body_SetAngularDamping( b2BodyId bodyId, float v )
{
	OR_BODY( bodyId );
	or_get_body_sim( world, body )->angularDamping = v;
}
B2_API float b2Body_GetAngularDamping( b2BodyId bodyId )
{
	OR_BODY( bodyId );
	return or_get_body_sim( world, body )->angularDamping;
}
B2_API void b2Body_SetGravityScale( b2BodyId bodyId, float v )
{
	OR_BODY( bodyId );
	or_get_body_sim( world, body )->gravityScale = v;
}
B2_API float b2Body_GetGravityScale( b2BodyId bodyId )
{
	OR_BODY( bodyId );
	return or_get_body_sim( world, body )->gravityScale;
}
B2_API bool b2Body_IsAwake( b2BodyId bodyId )
{
	OR_BODY( bodyId );
	return body->setIndex == 2;
}
B2_API bool b2Body_IsBullet( b2BodyId bodyId )
{
	OR_BODY( bodyId );
	return or_get_body_sim( world, body )->isBullet;
}
B2_API int b2Body_GetShapeCount( b2BodyId bodyId )
{
	OR_BODY( bodyId );
	return body->shapeCount;
}
B2_API int b2Body_GetShapes( b2BodyId bodyId, b2ShapeId* shapeArray, int capacity )
{
	OR_BODY( bodyId );
	int n = 0;
	for ( int id = body->headShapeId; id != OR_NULL && n < capacity; id = world->shapes.data[id].nextShapeId )
	{
		orShape* s = world->shapes.data + id;
		shapeArray[n++] = ( b2ShapeId ){ s->id + 1, bodyId.world0, s->revision };
	}
	return n;
}
B2_API int b2Body_GetContactCapacity( b2BodyId bodyId )
{
	OR_BODY( bodyId );
	return body->contactCount;
}

static orContactSim* or_get_contact_sim( orWorld* world, orContact* contact )
{
	if ( contact->colorIndex != OR_NULL )
		return world->colors[contact->colorIndex].contacts.data + contact->localIndex;
	return world->awakeContacts.data + contact->localIndex;
}

// src/body.c:447-489
B2_API int b2Body_GetContactData( b2BodyId bodyId, b2ContactData* contactData, int capacity )
{
	OR_BODY( bodyId );
	int index = 0;
	int key = body->headContactKey;
	while ( key != OR_NULL && index < capacity )
	{
		orContact* contact = world->contacts.data + ( key >> 1 );
		if ( contact->flags & OR_CONTACT_TOUCHING )
		{
			orShape* a = world->shapes.data + contact->shapeIdA;
			orShape* b = world->shapes.data + contact->shapeIdB;
			contactData[index].shapeIdA = ( b2ShapeId ){ a->id + 1, bodyId.world0, a->revision };
			contactData[index].shapeIdB = ( b2ShapeId ){ b->id + 1, bodyId.world0, b->revision };
			contactData[index].manifold = or_get_contact_sim( world, contact )->manifold;
			index += 1;
		}
		key = contact->edges[key & 1].nextKey;
	}
	return index;
}

B2_API b2AABB b2Body_ComputeAABB( b2BodyId bodyId )
{
	OR_BODY( bodyId );
	if ( body->headShapeId == OR_NULL )
	{
		b2Transform xf = or_get_body_sim( world, body )->transform;
		return ( b2AABB ){ xf.p, xf.p };
	}
	orShape* s = world->shapes.data + body->headShapeId;
	b2AABB aabb = s->aabb;
	while ( s->nextShapeId != OR_NULL )
	{
		s = world->shapes.data + s->nextShapeId;
		aabb = b2AABB_Union( aabb, s->aabb );
	}
	return aabb;
}

static void or_buffer_move( orWorld* world, int proxyKey )
{
	// src/broad_phase.h:76-83
	if ( or_bitset_get( &world->moveSet, (uint32_t)proxyKey ) == false )
	{
		or_bitset_set( &world->moveSet, (uint32_t)proxyKey );
		OR_PUSH( world->moveArray, proxyKey );
	}
}

// src/body.c:683-744
B2_API void b2Body_SetTransform( b2BodyId bodyId, b2Vec2 position, b2Rot rotation )
{
	OR_BODY( bodyId );
	OR_ASSERT( world->locked == false );
	orBodySim* sim = or_get_body_sim( world, body );
	sim->transform.p = position;
	sim->transform.q = rotation;
	sim->center = b2TransformPoint( sim->transform, sim->localCenter );
	sim->rotation0 = sim->transform.q;
	sim->center0 = sim->center;
#if ORACLE_SOLVER_F64
	sim->centerR = ( orVec2r ){ sim->center.x, sim->center.y };
	sim->qR = ( orVec2r ){ rotation.c, rotation.s };
#endif
	for ( int id = body->headShapeId; id != OR_NULL; id = world->shapes.data[id].nextShapeId )
	{
		orShape* shape = world->shapes.data + id;
		b2AABB aabb = or_compute_shape_aabb( shape, sim->transform );
		aabb.lowerBound.x -= OR_SPECULATIVE;
		aabb.lowerBound.y -= OR_SPECULATIVE;
		aabb.upperBound.x += OR_SPECULATIVE;
		aabb.upperBound.y += OR_SPECULATIVE;
		shape->aabb = aabb;
		if ( b2AABB_Contains( shape->fatAABB, aabb ) == false )
		{
			b2AABB fat = { { aabb.lowerBound.x - OR_AABB_MARGIN, aabb.lowerBound.y - OR_AABB_MARGIN },
						   { aabb.upperBound.x + OR_AABB_MARGIN, aabb.upperBound.y + OR_AABB_MARGIN } };
			shape->fatAABB = fat;
			int type = shape->proxyKey & 3;
			or_tree_move_proxy( world->trees + type, shape->proxyKey >> 2, fat );
			or_buffer_move( world, shape->proxyKey );
		}
	}
}

// ---- shapes (src/shape.c:38-217) ----
static b2Vec2 or_shape_centroid( const orShape* s )
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

static b2ShapeId or_create_shape( b2BodyId bodyId, const b2ShapeDef* def, const void* geometry, b2ShapeType type )
{
	OR_ASSERT( def->internalValue == OR_COOKIE );
	OR_ASSERT( def->isSensor == false && "sensor shapes are not supported" );
	orWorld* world = or_get_world_index( bodyId.world0 );
	if ( world->locked )
		return b2_nullShapeId;
	orBody* body = or_get_body_full( world, bodyId );
	b2Transform transform = or_get_body_sim( world, body )->transform;

	int shapeId = or_alloc_id( &world->shapeIdPool );
	if ( shapeId == world->shapes.count )
	{
		orShape empty = { 0 };
		OR_PUSH( world->shapes, empty );
	}
	orShape* shape = world->shapes.data + shapeId;
	uint16_t revision = shape->revision;
	memset( shape, 0, sizeof( *shape ) );
	switch ( type )
	{
		case b2_capsuleShape:
			shape->capsule = *(const b2Capsule*)geometry;
			break;
		case b2_circleShape:
			shape->circle = *(const b2Circle*)geometry;
			break;
		case b2_polygonShape:
			shape->polygon = *(const b2Polygon*)geometry;
			break;
		case b2_segmentShape:
			shape->segment = *(const b2Segment*)geometry;
			break;
		default:
			OR_ASSERT( false );
	}
	shape->id = shapeId;
	shape->bodyId = body->id;
	shape->type = type;
	shape->density = def->density;
	shape->friction = def->friction;
	shape->restitution = def->restitution;
	shape->filter = def->filter;
	shape->userData = def->userData;
	shape->isSensor = def->isSensor;
	shape->enableContactEvents = def->enableContactEvents;
	shape->enableHitEvents = def->enableHitEvents;
	shape->proxyKey = OR_NULL;
	shape->localCentroid = or_shape_centroid( shape );
	shape->revision = revision + 1;

	// b2CreateShapeProxy -> b2UpdateShapeAABBs (shape.c:38-59, 692-702)
	{
		b2AABB aabb = or_compute_shape_aabb( shape, transform );
		aabb.lowerBound.x -= OR_SPECULATIVE;
		aabb.lowerBound.y -= OR_SPECULATIVE;
		aabb.upperBound.x += OR_SPECULATIVE;
		aabb.upperBound.y += OR_SPECULATIVE;
		shape->aabb = aabb;
		float margin = body->type == b2_staticBody ? OR_SPECULATIVE : OR_AABB_MARGIN;
		shape->fatAABB = ( b2AABB ){ { aabb.lowerBound.x - margin, aabb.lowerBound.y - margin },
									 { aabb.upperBound.x + margin, aabb.upperBound.y + margin } };
		int proxyType = body->type;
		int proxyId = or_tree_create_proxy( world->trees + proxyType, shape->fatAABB, shape->filter.categoryBits, shapeId );
		shape->proxyKey = ( proxyId << 2 ) | proxyType;
		if ( proxyType != b2_staticBody || def->forceContactCreation )
			or_buffer_move( world, shape->proxyKey );
	}

	// head insertion into the body's shape list (shape.c:137-148)
	if ( body->headShapeId != OR_NULL )
		world->shapes.data[body->headShapeId].prevShapeId = shapeId;
	shape->prevShapeId = OR_NULL;
	shape->nextShapeId = body->headShapeId;
	body->headShapeId = shapeId;
	body->shapeCount += 1;

	if ( body->automaticMass )
		or_update_body_mass( world, body );
	return ( b2ShapeId ){ shapeId + 1, bodyId.world0, shape->revision };
}

B2_API b2ShapeId b2CreateCircleShape( b2BodyId bodyId, const b2ShapeDef* def, const b2Circle* circle )
{
	return or_create_shape( bodyId, def, circle, b2_circleShape );
}

B2_API b2ShapeId b2CreateCapsuleShape( b2BodyId bodyId, const b2ShapeDef* def, const b2Capsule* capsule )
{
	float lengthSqr = b2DistanceSquared( capsule->center1, capsule->center2 );
	if ( lengthSqr <= OR_LINEAR_SLOP * OR_LINEAR_SLOP )
	{
		b2Circle circle = { b2Lerp( capsule->center1, capsule->center2, 0.5f ), capsule->radius };
		return or_create_shape( bodyId, def, &circle, b2_circleShape );
	}
	return or_create_shape( bodyId, def, capsule, b2_capsuleShape );
}

B2_API b2ShapeId b2CreatePolygonShape( b2BodyId bodyId, const b2ShapeDef* def, const b2Polygon* polygon )
{
	return or_create_shape( bodyId, def, polygon, b2_polygonShape );
}

B2_API b2ShapeId b2CreateSegmentShape( b2BodyId bodyId, const b2ShapeDef* def, const b2Segment* segment )
{
	float lengthSqr = b2DistanceSquared( segment->point1, segment->point2 );
	if ( lengthSqr <= OR_LINEAR_SLOP * OR_LINEAR_SLOP )
		return b2_nullShapeId;
	return or_create_shape( bodyId, def, segment, b2_segmentShape );
}

B2_API bool b2Shape_IsValid( b2ShapeId id )
{
	if ( id.world0 >= OR_MAX_WORLDS )
		return false;
	orWorld* world = or_worlds + id.world0;
	if ( world->inUse == false || id.index1 < 1 || world->shapes.count < id.index1 )
		return false;
	orShape* s = world->shapes.data + ( id.index1 - 1 );
	return s->id == id.index1 - 1 && s->revision == id.revision;
}

#define OR_SHAPE( shapeId ) \
	orWorld* world = or_get_world_index( ( shapeId ).world0 ); \
	orShape* shape = or_get_shape_full( world, ( shapeId ) )

B2_API b2ShapeType b2Shape_GetType( b2ShapeId shapeId )
{
	OR_SHAPE( shapeId );
	return shape->type;
}
B2_API b2BodyId b2Shape_GetBody( b2ShapeId shapeId )
{
	OR_SHAPE( shapeId );
	orBody* body = world->bodies.data + shape->bodyId;
	return ( b2BodyId ){ body->id + 1, shapeId.world0, body->revision };
}
B2_API float b2Shape_GetDensity( b2ShapeId shapeId )
{
	OR_SHAPE( shapeId );
	return shape->density;
}
B2_API void b2Shape_SetFriction( b2ShapeId shapeId, float friction )
{
	OR_SHAPE( shapeId );
	shape->friction = friction;
}
B2_API float b2Shape_GetFriction( b2ShapeId shapeId )
{
	OR_SHAPE( shapeId );
	return shape->friction;
}
B2_API void b2Shape_SetRestitution( b2ShapeId shapeId, float restitution )
{
	OR_SHAPE( shapeId );
	shape->restitution = restitution;
}
B2_API float b2Shape_GetRestitution( b2ShapeId shapeId )
{
	OR_SHAPE( shapeId );
	return shape->restitution;
}
B2_API b2Filter b2Shape_GetFilter( b2ShapeId shapeId )
{
	OR_SHAPE( shapeId );
	return shape->filter;
}
B2_API b2Circle b2Shape_GetCircle( b2ShapeId shapeId )
{
	OR_SHAPE( shapeId );
	return shape->circle;
}
B2_API b2Segment b2Shape_GetSegment( b2ShapeId shapeId )
{
	OR_SHAPE( shapeId );
	return shape->segment;
}
B2_API b2Capsule b2Shape_GetCapsule( b2ShapeId shapeId )
{
	OR_SHAPE( shapeId );
	return shape->capsule;
}
B2_API b2Polygon b2Shape_GetPolygon( b2ShapeId shapeId )
{
	OR_SHAPE( shapeId );
	return shape->polygon;
}
B2_API b2AABB b2Shape_GetAABB( b2ShapeId shapeId )
{
	OR_SHAPE( shapeId );
	return shape->aabb;
}
B2_API int b2Shape_GetContactCapacity( b2ShapeId shapeId )
{
	OR_SHAPE( shapeId );
	return world->bodies.data[shape->bodyId].contactCount;
}
// src/shape.c:1102-1140
B2_API int b2Shape_GetContactData( b2ShapeId shapeId, b2ContactData* contactData, int capacity )
{
	OR_SHAPE( shapeId );
	orBody* body = world->bodies.data + shape->bodyId;
	int index = 0;
	int key = body->headContactKey;
	while ( key != OR_NULL && index < capacity )
	{
		orContact* contact = world->contacts.data + ( key >> 1 );
		if ( ( contact->shapeIdA == shape->id || contact->shapeIdB == shape->id ) && ( contact->flags & OR_CONTACT_TOUCHING ) )
		{
			orShape* a = world->shapes.data + contact->shapeIdA;
			orShape* b = world->shapes.data + contact->shapeIdB;
			contactData[index].shapeIdA = ( b2ShapeId ){ a->id + 1, shapeId.world0, a->revision };
			contactData[index].shapeIdB = ( b2ShapeId ){ b->id + 1, shapeId.world0, b->revision };
			contactData[index].manifold = or_get_contact_sim( world, contact )->manifold;
			index += 1;
		}
		key = contact->edges[key & 1].nextKey;
	}
	return index;
}

// ---- contacts (src/contact.c:185-414) ----
void or_create_contact( orWorld* world, orShape* shapeA, orShape* shapeB )
{
	bool primary;
	if ( or_has_manifold_fcn( shapeA->type, shapeB->type, &primary ) == false )
		return;
	if ( primary == false )
	{
		or_create_contact( world, shapeB, shapeA );
		return;
	}
	orBody* bodyA = world->bodies.data + shapeA->bodyId;
	orBody* bodyB = world->bodies.data + shapeB->bodyId;
	OR_ASSERT( bodyA->setIndex == 2 || bodyB->setIndex == 2 );

	int contactId = or_alloc_id( &world->contactIdPool );
	if ( contactId == world->contacts.count )
	{
		orContact empty = { 0 };
		OR_PUSH( world->contacts, empty );
	}
	orContact* contact = world->contacts.data + contactId;
	contact->contactId = contactId;
	contact->setIndex = 2;
	contact->colorIndex = OR_NULL;
	contact->localIndex = world->awakeContacts.count;
	contact->shapeIdA = shapeA->id;
	contact->shapeIdB = shapeB->id;
	contact->flags = 0;
	if ( shapeA->enableContactEvents || shapeB->enableContactEvents )
		contact->flags |= OR_CONTACT_ENABLE_CONTACT_EVENTS;

	// connect to body A then body B (head insertion of edge keys)
	for ( int e = 0; e < 2; ++e )
	{
		orBody* body = e == 0 ? bodyA : bodyB;
		contact->edges[e].bodyId = body->id;
		contact->edges[e].prevKey = OR_NULL;
		contact->edges[e].nextKey = body->headContactKey;
		int key = ( contactId << 1 ) | e;
		int head = body->headContactKey;
		if ( head != OR_NULL )
			world->contacts.data[head >> 1].edges[head & 1].prevKey = key;
		body->headContactKey = key;
		body->contactCount += 1;
	}

	or_keyset_add( &world->pairSet, or_pair_key( shapeA->id, shapeB->id ) );

	orContactSim sim = { 0 };
	sim.contactId = contactId;
	sim.bodySimIndexA = OR_NULL;
	sim.bodySimIndexB = OR_NULL;
	sim.shapeIdA = shapeA->id;
	sim.shapeIdB = shapeB->id;
	sim.friction = sqrtf( shapeA->friction * shapeB->friction );
	sim.restitution = shapeA->restitution > shapeB->restitution ? shapeA->restitution : shapeB->restitution;
	OR_PUSH( world->awakeContacts, sim );
}

// src/constraint_graph.c:64-183
void or_add_contact_to_graph( orWorld* world, orContactSim* contactSim, orContact* contact )
{
	int colorIndex = OR_OVERFLOW;
	int bodyIdA = contact->edges[0].bodyId;
	int bodyIdB = contact->edges[1].bodyId;
	orBody* bodyA = world->bodies.data + bodyIdA;
	orBody* bodyB = world->bodies.data + bodyIdB;
	bool staticA = bodyA->setIndex == 0;
	bool staticB = bodyB->setIndex == 0;

	if ( staticA == false && staticB == false )
	{
		for ( int i = 0; i < OR_OVERFLOW; ++i )
		{
			orBitSet* set = &world->colors[i].bodySet;
			if ( or_bitset_get( set, bodyIdA ) || or_bitset_get( set, bodyIdB ) )
				continue;
			or_bitset_set( set, bodyIdA );
			or_bitset_set( set, bodyIdB );
			colorIndex = i;
			break;
		}
	}
	else if ( staticA == false )
	{
		for ( int i = 1; i < OR_OVERFLOW; ++i )
		{
			orBitSet* set = &world->colors[i].bodySet;
			if ( or_bitset_get( set, bodyIdA ) )
				continue;
			or_bitset_set( set, bodyIdA );
			colorIndex = i;
			break;
		}
	}
	else if ( staticB == false )
	{
		for ( int i = 1; i < OR_OVERFLOW; ++i )
		{
			orBitSet* set = &world->colors[i].bodySet;
			if ( or_bitset_get( set, bodyIdB ) )
				continue;
			or_bitset_set( set, bodyIdB );
			colorIndex = i;
			break;
		}
	}

	orGraphColor* color = world->colors + colorIndex;
	contact->colorIndex = colorIndex;
	contact->localIndex = color->contacts.count;
	orContactSim copy = *contactSim;
	if ( staticA )
	{
		copy.bodySimIndexA = OR_NULL;
		copy.invMassA = 0.0f;
		copy.invIA = 0.0f;
	}
	else
	{
		copy.bodySimIndexA = bodyA->localIndex;
		copy.invMassA = world->awakeSims.data[bodyA->localIndex].invMass;
		copy.invIA = world->awakeSims.data[bodyA->localIndex].invInertia;
	}
	if ( staticB )
	{
		copy.bodySimIndexB = OR_NULL;
		copy.invMassB = 0.0f;
		copy.invIB = 0.0f;
	}
	else
	{
		copy.bodySimIndexB = bodyB->localIndex;
		copy.invMassB = world->awakeSims.data[bodyB->localIndex].invMass;
		copy.invIB = world->awakeSims.data[bodyB->localIndex].invInertia;
	}
	OR_PUSH( color->contacts, copy );
}

// src/constraint_graph.c:185-214
void or_remove_contact_from_graph( orWorld* world, int bodyIdA, int bodyIdB, int colorIndex, int localIndex )
{
	orGraphColor* color = world->colors + colorIndex;
	if ( colorIndex != OR_OVERFLOW )
	{
		or_bitset_clear( &color->bodySet, bodyIdA );
		or_bitset_clear( &color->bodySet, bodyIdB );
	}
	int last = color->contacts.count - 1;
	if ( localIndex < last )
	{
		color->contacts.data[localIndex] = color->contacts.data[last];
		int movedId = color->contacts.data[localIndex].contactId;
		world->contacts.data[movedId].localIndex = localIndex;
	}
	color->contacts.count -= 1;
}

static void or_remove_non_touching( orWorld* world, int localIndex )
{
	int last = world->awakeContacts.count - 1;
	if ( localIndex < last )
	{
		world->awakeContacts.data[localIndex] = world->awakeContacts.data[last];
		int movedId = world->awakeContacts.data[localIndex].contactId;
		world->contacts.data[movedId].localIndex = localIndex;
	}
	world->awakeContacts.count -= 1;
}

// src/contact.c:335-414
void or_destroy_contact( orWorld* world, orContact* contact )
{
	or_keyset_remove( &world->pairSet, or_pair_key( contact->shapeIdA, contact->shapeIdB ) );
	int contactId = contact->contactId;
	for ( int e = 0; e < 2; ++e )
	{
		orContactEdge* edge = contact->edges + e;
		orBody* body = world->bodies.data + edge->bodyId;
		if ( edge->prevKey != OR_NULL )
			world->contacts.data[edge->prevKey >> 1].edges[edge->prevKey & 1].nextKey = edge->nextKey;
		if ( edge->nextKey != OR_NULL )
			world->contacts.data[edge->nextKey >> 1].edges[edge->nextKey & 1].prevKey = edge->prevKey;
		int key = ( contactId << 1 ) | e;
		if ( body->headContactKey == key )
			body->headContactKey = edge->nextKey;
		body->contactCount -= 1;
	}
	if ( contact->colorIndex != OR_NULL )
		or_remove_contact_from_graph( world, contact->edges[0].bodyId, contact->edges[1].bodyId, contact->colorIndex,
									  contact->localIndex );
	else
		or_remove_non_touching( world, contact->localIndex );
	contact->contactId = OR_NULL;
	contact->setIndex = OR_NULL;
	contact->colorIndex = OR_NULL;
	contact->localIndex = OR_NULL;
	or_free_id( &world->contactIdPool, contactId );
}

// ---- destroy body (src/body.c:320-440 subset: contacts, shapes, proxies, sims) ----
B2_API void b2DestroyBody( b2BodyId bodyId )
{
	OR_BODY( bodyId );
	OR_ASSERT( world->locked == false );
	// the attached joints first (src/body.c:349-362)
	for ( int jointKey = body->headJointKey; jointKey != OR_NULL; )
	{
		orJoint* joint = world->joints.data + ( jointKey >> 1 );
		jointKey = joint->edges[jointKey & 1].nextKey;
		or_destroy_joint_internal( world, joint );
	}
	int key = body->headContactKey;
	while ( key != OR_NULL )
	{
		orContact* contact = world->contacts.data + ( key >> 1 );
		key = contact->edges[key & 1].nextKey;
		or_destroy_contact( world, contact );
	}
	for ( int id = body->headShapeId; id != OR_NULL; )
	{
		orShape* shape = world->shapes.data + id;
		id = shape->nextShapeId;
		// unbuffer move (broad_phase.c:71-89) and destroy proxy
		if ( or_bitset_get( &world->moveSet, (uint32_t)shape->proxyKey ) )
		{
			or_bitset_clear( &world->moveSet, (uint32_t)shape->proxyKey );
			for ( int i = 0; i < world->moveArray.count; ++i )
			{
				if ( world->moveArray.data[i] == shape->proxyKey )
				{
					world->moveArray.data[i] = world->moveArray.data[world->moveArray.count - 1];
					world->moveArray.count -= 1;
					break;
				}
			}
		}
		or_tree_destroy_proxy( world->trees + ( shape->proxyKey & 3 ), shape->proxyKey >> 2 );
		or_free_id( &world->shapeIdPool, shape->id );
		shape->id = OR_NULL;
	}
	// swap-remove the sim (and state) from its set
	if ( body->setIndex == 0 )
	{
		int last = world->staticSims.count - 1;
		if ( body->localIndex < last )
		{
			world->staticSims.data[body->localIndex] = world->staticSims.data[last];
			world->bodies.data[world->staticSims.data[body->localIndex].bodyId].localIndex = body->localIndex;
		}
		world->staticSims.count -= 1;
	}
	else
	{
		int last = world->awakeSims.count - 1;
		if ( body->localIndex < last )
		{
			world->awakeSims.data[body->localIndex] = world->awakeSims.data[last];
			world->awakeStates.data[body->localIndex] = world->awakeStates.data[last];
			world->bodies.data[world->awakeSims.data[body->localIndex].bodyId].localIndex = body->localIndex;
		}
		world->awakeSims.count -= 1;
		world->awakeStates.count -= 1;
	}
	or_free_id( &world->bodyIdPool, body->id );
	body->setIndex = OR_NULL;
	body->localIndex = OR_NULL;
	body->id = OR_NULL;
}

// ---- broadphase pair update (src/broad_phase.c:162-436) ----
typedef struct orMovePair
{
	int shapeIndexA, shapeIndexB;
	int next;
} orMovePair;

typedef OR_ARRAY( orMovePair ) orMovePairArray;

typedef struct orPairQuery
{
	orWorld* world;
	orMovePairArray* pairs;
	int* head;
	int queryTreeType;
	int queryProxyKey;
	int queryShapeIndex;
} orPairQuery;

static bool or_should_shapes_collide( b2Filter a, b2Filter b )
{
	// src/contact.c:465-474
	if ( a.groupIndex == b.groupIndex && a.groupIndex != 0 )
		return a.groupIndex > 0;
	return ( a.maskBits & b.categoryBits ) != 0 && ( a.categoryBits & b.maskBits ) != 0;
}

static bool or_pair_query_callback( int proxyId, int shapeId, void* context )
{
	orPairQuery* q = context;
	orWorld* world = q->world;
	int proxyKey = ( proxyId << 2 ) | q->queryTreeType;
	if ( proxyKey == q->queryProxyKey )
		return true;
	if ( q->queryTreeType != b2_staticBody )
	{
		bool moved = or_bitset_get( &world->moveSet, (uint32_t)proxyKey );
		if ( moved && proxyKey < q->queryProxyKey )
			return true;
	}
	if ( or_keyset_contains( &world->pairSet, or_pair_key( shapeId, q->queryShapeIndex ) ) )
		return true;

	int shapeIdA, shapeIdB;
	if ( proxyKey < q->queryProxyKey )
	{
		shapeIdA = shapeId;
		shapeIdB = q->queryShapeIndex;
	}
	else
	{
		shapeIdA = q->queryShapeIndex;
		shapeIdB = shapeId;
	}
	orShape* shapeA = world->shapes.data + shapeIdA;
	orShape* shapeB = world->shapes.data + shapeIdB;
	if ( shapeA->bodyId == shapeB->bodyId )
		return true;
	if ( or_should_shapes_collide( shapeA->filter, shapeB->filter ) == false )
		return true;
	// b2ShouldBodiesCollide (body.c:1723-1760): at least one dynamic body, no joint between them
	// that turns their collision off
	orBody* bodyA = world->bodies.data + shapeA->bodyId;
	orBody* bodyB = world->bodies.data + shapeB->bodyId;
	if ( or_should_bodies_collide( world, bodyA, bodyB ) == false )
		return true;

	orMovePair pair = { shapeIdA, shapeIdB, *q->head };
	*q->head = q->pairs->count;
	OR_PUSH( *q->pairs, pair );
	return true;
}

void or_update_pairs( orWorld* world )
{
	int moveCount = world->moveArray.count;
	if ( moveCount == 0 )
		return;

	orMovePairArray pairs = { 0 };
	int* heads = malloc( moveCount * sizeof( int ) );
	orPairQuery q;
	q.world = world;
	q.pairs = &pairs;
	for ( int i = 0; i < moveCount; ++i )
	{
		heads[i] = OR_NULL;
		q.head = heads + i;
		int proxyKey = world->moveArray.data[i];
		int proxyType = proxyKey & 3;
		int proxyId = proxyKey >> 2;
		q.queryProxyKey = proxyKey;
		const orTree* baseTree = world->trees + proxyType;
		b2AABB fatAABB = baseTree->nodes[proxyId].aabb;
		q.queryShapeIndex = baseTree->nodes[proxyId].userData;
		if ( proxyType == b2_dynamicBody )
		{
			q.queryTreeType = b2_kinematicBody;
			or_tree_query( world->trees + b2_kinematicBody, fatAABB, b2_defaultMaskBits, or_pair_query_callback, &q );
			q.queryTreeType = b2_staticBody;
			or_tree_query( world->trees + b2_staticBody, fatAABB, b2_defaultMaskBits, or_pair_query_callback, &q );
		}
		q.queryTreeType = b2_dynamicBody;
		or_tree_query( world->trees + b2_dynamicBody, fatAABB, b2_defaultMaskBits, or_pair_query_callback, &q );
	}

	// create contacts in move-array order, each list head first (broad_phase.c:378-415)
	for ( int i = 0; i < moveCount; ++i )
	{
		for ( int p = heads[i]; p != OR_NULL; p = pairs.data[p].next )
			or_create_contact( world, world->shapes.data + pairs.data[p].shapeIndexA, world->shapes.data + pairs.data[p].shapeIndexB );
	}
	free( heads );
	OR_FREE( pairs );

	// reset the move buffer
	for ( int i = 0; i < moveCount; ++i )
		or_bitset_clear( &world->moveSet, (uint32_t)world->moveArray.data[i] );
	world->moveArray.count = 0;
}

// ---- narrowphase driver (src/world.c:290-674, src/contact.c:493-582) ----
static const void* or_shape_geometry( const orShape* s )
{
	return &s->circle; // union: every member starts at the same address
}

static bool or_update_contact( orContactSim* cs, orShape* shapeA, b2Transform xfA, b2Vec2 centerOffsetA, orShape* shapeB,
							   b2Transform xfB, b2Vec2 centerOffsetB )
{
	b2Manifold oldManifold = cs->manifold;
	cs->manifold = or_collide_shapes( shapeA->type, or_shape_geometry( shapeA ), xfA, shapeB->type, or_shape_geometry( shapeB ), xfB );
	int pointCount = cs->manifold.pointCount;
	bool touching = pointCount > 0;
	if ( touching && ( shapeA->enableHitEvents || shapeB->enableHitEvents ) )
		cs->simFlags |= OR_SIM_HIT_EVENT;
	else
		cs->simFlags &= ~OR_SIM_HIT_EVENT;

	for ( int i = 0; i < pointCount; ++i )
	{
		b2ManifoldPoint* mp2 = cs->manifold.points + i;
		mp2->anchorA = b2Sub( mp2->anchorA, centerOffsetA );
		mp2->anchorB = b2Sub( mp2->anchorB, centerOffsetB );
		mp2->normalImpulse = 0.0f;
		mp2->tangentImpulse = 0.0f;
		mp2->maxNormalImpulse = 0.0f;
		mp2->normalVelocity = 0.0f;
		mp2->persisted = false;
		uint16_t id2 = mp2->id;
		for ( int j = 0; j < oldManifold.pointCount; ++j )
		{
			b2ManifoldPoint* mp1 = oldManifold.points + j;
			if ( mp1->id == id2 )
			{
				mp2->normalImpulse = mp1->normalImpulse;
				mp2->tangentImpulse = mp1->tangentImpulse;
				mp2->persisted = true;
				break;
			}
		}
	}
	if ( touching )
		cs->simFlags |= OR_SIM_TOUCHING;
	else
		cs->simFlags &= ~OR_SIM_TOUCHING;
	return touching;
}

static void or_collide_one( orWorld* world, orContactSim* cs, orBitSet* stateBits )
{
	orShape* shapeA = world->shapes.data + cs->shapeIdA;
	orShape* shapeB = world->shapes.data + cs->shapeIdB;
	if ( or_aabb_overlaps( shapeA->fatAABB, shapeB->fatAABB ) == false )
	{
		cs->simFlags |= OR_SIM_DISJOINT;
		cs->simFlags &= ~OR_SIM_TOUCHING;
		or_bitset_set( stateBits, cs->contactId );
		return;
	}
	bool wasTouching = ( cs->simFlags & OR_SIM_TOUCHING ) != 0;
	orBody* bodyA = world->bodies.data + shapeA->bodyId;
	orBody* bodyB = world->bodies.data + shapeB->bodyId;
	orBodySim* simA = or_get_body_sim( world, bodyA );
	orBodySim* simB = or_get_body_sim( world, bodyB );
	cs->bodySimIndexA = bodyA->setIndex == 2 ? bodyA->localIndex : OR_NULL;
	cs->invMassA = simA->invMass;
	cs->invIA = simA->invInertia;
	cs->bodySimIndexB = bodyB->setIndex == 2 ? bodyB->localIndex : OR_NULL;
	cs->invMassB = simB->invMass;
	cs->invIB = simB->invInertia;
	b2Transform xfA = simA->transform, xfB = simB->transform;
	b2Vec2 offsetA = b2RotateVector( xfA.q, simA->localCenter );
	b2Vec2 offsetB = b2RotateVector( xfB.q, simB->localCenter );
	bool touching = or_update_contact( cs, shapeA, xfA, offsetA, shapeB, xfB, offsetB );
	if ( touching && !wasTouching )
	{
		cs->simFlags |= OR_SIM_STARTED;
		or_bitset_set( stateBits, cs->contactId );
	}
	else if ( !touching && wasTouching )
	{
		cs->simFlags |= OR_SIM_STOPPED;
		or_bitset_set( stateBits, cs->contactId );
	}
}

void or_collide( orWorld* world, orStepContext* ctx )
{
	(void)ctx;
	// b2UpdateTreesTask (world.c:367-379), enqueued first in b2Collide (world.c:420)
	or_tree_rebuild( world->trees + b2_dynamicBody, false );
	or_tree_rebuild( world->trees + b2_kinematicBody, false );

	orBitSet stateBits = { 0 };
	or_bitset_set_count_and_clear( &stateBits, (uint32_t)world->contactIdPool.nextIndex );
	for ( int c = 0; c < OR_GRAPH_COLORS; ++c )
	{
		orGraphColor* color = world->colors + c;
		for ( int i = 0; i < color->contacts.count; ++i )
			or_collide_one( world, color->contacts.data + i, &stateBits );
	}
	for ( int i = 0; i < world->awakeContacts.count; ++i )
		or_collide_one( world, world->awakeContacts.data + i, &stateBits );

	// serial state changes in contact-id order (world.c:506-667), with their begin/end touch
	// events (world.c:544-551, 583-590, 632-639)
	for ( uint32_t k = 0; k < stateBits.blockCount; ++k )
	{
		uint64_t bits = stateBits.bits[k];
		while ( bits != 0 )
		{
			int contactId = (int)( 64 * k + __builtin_ctzll( bits ) );
			orContact* contact = world->contacts.data + contactId;
			int colorIndex = contact->colorIndex;
			int localIndex = contact->localIndex;
			orContactSim* cs = or_get_contact_sim( world, contact );
			uint32_t simFlags = cs->simFlags;
			const orShape* shapeA = world->shapes.data + contact->shapeIdA;
			const orShape* shapeB = world->shapes.data + contact->shapeIdB;
			b2ShapeId shapeIdA = { shapeA->id + 1, (uint16_t)world->worldId, shapeA->revision };
			b2ShapeId shapeIdB = { shapeB->id + 1, (uint16_t)world->worldId, shapeB->revision };
			bool events = ( contact->flags & OR_CONTACT_ENABLE_CONTACT_EVENTS ) != 0;
			if ( simFlags & OR_SIM_DISJOINT )
			{
				if ( ( contact->flags & OR_CONTACT_TOUCHING ) && events )
				{
					b2ContactEndTouchEvent e = { shapeIdA, shapeIdB };
					OR_PUSH( world->contactEnd, e );
				}
				contact->flags &= ~OR_CONTACT_TOUCHING;
				or_destroy_contact( world, contact );
			}
			else if ( simFlags & OR_SIM_STARTED )
			{
				if ( events )
				{
					b2ContactBeginTouchEvent e = { shapeIdA, shapeIdB };
					OR_PUSH( world->contactBegin, e );
				}
				contact->flags |= OR_CONTACT_TOUCHING;
				cs->simFlags &= ~OR_SIM_STARTED;
				or_add_contact_to_graph( world, cs, contact );
				or_remove_non_touching( world, localIndex );
			}
			else if ( simFlags & OR_SIM_STOPPED )
			{
				cs->simFlags &= ~OR_SIM_STOPPED;
				contact->flags &= ~OR_CONTACT_TOUCHING;
				if ( contact->flags & OR_CONTACT_ENABLE_CONTACT_EVENTS )
				{
					b2ContactEndTouchEvent e = { shapeIdA, shapeIdB };
					OR_PUSH( world->contactEnd, e );
				}
				int bodyIdA = contact->edges[0].bodyId, bodyIdB = contact->edges[1].bodyId;
				// b2AddNonTouchingContact then b2RemoveContactFromGraph
				contact->colorIndex = OR_NULL;
				contact->localIndex = world->awakeContacts.count;
				orContactSim copy = *cs;
				OR_PUSH( world->awakeContacts, copy );
				or_remove_contact_from_graph( world, bodyIdA, bodyIdB, colorIndex, localIndex );
			}
			bits &= bits - 1;
		}
	}
	or_bitset_free( &stateBits );
}

// ---- events (src/world.c:1366-1413; src/shape.c:1010-1065; src/body.c:1654-1664) ----
B2_API b2BodyEvents b2World_GetBodyEvents( b2WorldId worldId )
{
	orWorld* world = or_get_world( worldId );
	if ( world->locked )
		return ( b2BodyEvents ){ 0 };
	return ( b2BodyEvents ){ world->moveEvents.data, world->moveEvents.count };
}

B2_API b2SensorEvents b2World_GetSensorEvents( b2WorldId worldId )
{
	orWorld* world = or_get_world( worldId );
	if ( world->locked )
		return ( b2SensorEvents ){ 0 };
	return ( b2SensorEvents ){ world->sensorBegin.data, world->sensorEnd.data, world->sensorBegin.count, world->sensorEnd.count };
}

B2_API b2ContactEvents b2World_GetContactEvents( b2WorldId worldId )
{
	orWorld* world = or_get_world( worldId );
	if ( world->locked )
		return ( b2ContactEvents ){ 0 };
	return ( b2ContactEvents ){ world->contactBegin.data, world->contactEnd.data, world->contactHit.data,
								world->contactBegin.count, world->contactEnd.count, world->contactHit.count };
}

B2_API void b2Shape_EnableContactEvents( b2ShapeId shapeId, bool flag )
{
	orWorld* world = or_get_world_index( shapeId.world0 );
	if ( world->locked )
		return;
	or_get_shape_full( world, shapeId )->enableContactEvents = flag;
}

B2_API bool b2Shape_AreContactEventsEnabled( b2ShapeId shapeId )
{
	return or_get_shape_full( or_get_world_index( shapeId.world0 ), shapeId )->enableContactEvents;
}

B2_API void b2Shape_EnableHitEvents( b2ShapeId shapeId, bool flag )
{
	orWorld* world = or_get_world_index( shapeId.world0 );
	if ( world->locked )
		return;
	or_get_shape_full( world, shapeId )->enableHitEvents = flag;
}

B2_API bool b2Shape_AreHitEventsEnabled( b2ShapeId shapeId )
{
	return or_get_shape_full( or_get_world_index( shapeId.world0 ), shapeId )->enableHitEvents;
}

B2_API void b2Body_EnableHitEvents( b2BodyId bodyId, bool enableHitEvents )
{
	orWorld* world = or_get_world_index( bodyId.world0 );
	orBody* body = or_get_body_full( world, bodyId );
	for ( int id = body->headShapeId; id != OR_NULL; id = world->shapes.data[id].nextShapeId )
		world->shapes.data[id].enableHitEvents = enableHitEvents;
}

// ---- inspection (include/b2mi.h) ----
static void or_fill_record( b2mi_ContactRecord* r, const orContactSim* cs, int color, int local )
{
	r->contactId = cs->contactId;
	r->shapeIdA = cs->shapeIdA;
	r->shapeIdB = cs->shapeIdB;
	r->colorIndex = color;
	r->localIndex = local;
	r->simFlags = cs->simFlags;
	r->friction = cs->friction;
	r->restitution = cs->restitution;
	r->manifold = cs->manifold;
}

B2_API int b2mi_World_DumpContacts( b2WorldId worldId, b2mi_ContactRecord* records, int capacity )
{
	orWorld* world = or_get_world( worldId );
	int n = 0;
	for ( int c = 0; c < OR_GRAPH_COLORS; ++c )
		for ( int i = 0; i < world->colors[c].contacts.count; ++i, ++n )
			if ( n < capacity )
				or_fill_record( records + n, world->colors[c].contacts.data + i, c, i );
	for ( int i = 0; i < world->awakeContacts.count; ++i, ++n )
		if ( n < capacity )
			or_fill_record( records + n, world->awakeContacts.data + i, -1, i );
	return n;
}

B2_API int b2mi_World_DumpBodies( b2WorldId worldId, b2mi_BodyRecord* records, int capacity )
{
	orWorld* world = or_get_world( worldId );
	int n = 0;
	for ( int id = 0; id < world->bodies.count; ++id )
	{
		orBody* body = world->bodies.data + id;
		if ( body->setIndex == OR_NULL )
			continue;
		if ( n < capacity )
		{
			b2mi_BodyRecord* r = records + n;
			memset( r, 0, sizeof( *r ) );
			orBodySim* sim = or_get_body_sim( world, body );
			r->bodyId = id;
			r->type = body->type;
			r->setIndex = body->setIndex;
			r->localIndex = body->localIndex;
			r->transform = sim->transform;
			r->center = sim->center;
			orBodyState* st = or_get_body_state( world, body );
			if ( st )
			{
				r->linearVelocity = ( b2Vec2 ){ (float)st->v.x, (float)st->v.y };
				r->angularVelocity = (float)st->w;
			}
		}
		n += 1;
	}
	return n;
}

// The world's time-of-impact events so far (a probe for the tests that continuous collision moved
// some body back; not part of the reference's API).
B2_API int oracle_World_GetToiHits( b2WorldId worldId )
{
	return or_get_world( worldId )->toiHits;
}
