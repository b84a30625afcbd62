// Public C API of the MI355X step pipeline: the world / body / shape / step subset of the
// reference's include/box2d/box2d.h. Each declaration keeps the reference's name, signature
// and argument meaning; the reference line that declares it is cited so a maintainer can
// check the binding. In a step, the narrowphase, the solver stages and finalize run as HIP
// kernels on the GPU, and so do the broadphase pair queries of a world stepped alone with many
// moved proxies; the order-defining bookkeeping (dynamic-tree maintenance, contact creation,
// begin/end-touch processing and its events, graph colouring, work-item building) runs on the
// host. See b2mi.h for the device/batch/adjoint extensions.
#pragma once

#include "base.h"
#include "collision.h"
#include "id.h"
#include "types.h"

// ---- world (reference include/box2d/box2d.h:27-158) ----
B2_API b2WorldId b2CreateWorld( const b2WorldDef* def );											  // :27
B2_API void b2DestroyWorld( b2WorldId worldId );														  // :30
B2_API bool b2World_IsValid( b2WorldId id );															  // :33
B2_API void b2World_Step( b2WorldId worldId, float timeStep, int subStepCount );						  // :39
B2_API b2BodyEvents b2World_GetBodyEvents( b2WorldId worldId );											  // :45
B2_API b2SensorEvents b2World_GetSensorEvents( b2WorldId worldId );										  // :48
B2_API b2ContactEvents b2World_GetContactEvents( b2WorldId worldId );									  // :51
B2_API void b2World_EnableSleeping( b2WorldId worldId, bool flag );										  // :100
B2_API void b2World_EnableContinuous( b2WorldId worldId, bool flag );									  // :106
B2_API void b2World_SetRestitutionThreshold( b2WorldId worldId, float value );							  // :111
B2_API void b2World_SetHitEventThreshold( b2WorldId worldId, float value );								  // :116
B2_API void b2World_SetCustomFilterCallback( b2WorldId worldId, b2CustomFilterFcn* fcn, void* context ); // :119
B2_API void b2World_SetPreSolveCallback( b2WorldId worldId, b2PreSolveFcn* fcn, void* context );		  // :122
B2_API void b2World_SetGravity( b2WorldId worldId, b2Vec2 gravity );									  // :127
B2_API b2Vec2 b2World_GetGravity( b2WorldId worldId );													  // :130
B2_API void b2World_SetContactTuning( b2WorldId worldId, float hertz, float dampingRatio, float pushVelocity ); // :145
B2_API void b2World_EnableWarmStarting( b2WorldId worldId, bool flag );									  // :149
B2_API b2Profile b2World_GetProfile( b2WorldId worldId );												  // :152
B2_API b2Counters b2World_GetCounters( b2WorldId worldId );												  // :155

// ---- bodies (reference include/box2d/box2d.h:175-409) ----
B2_API b2BodyId b2CreateBody( b2WorldId worldId, const b2BodyDef* def );			 // :175
B2_API void b2DestroyBody( b2BodyId bodyId );										 // :179
B2_API bool b2Body_IsValid( b2BodyId id );											 // :182
B2_API b2BodyType b2Body_GetType( b2BodyId bodyId );								 // :185
B2_API void b2Body_SetUserData( b2BodyId bodyId, void* userData );					 // :192
B2_API void* b2Body_GetUserData( b2BodyId bodyId );								 // :195
B2_API b2Vec2 b2Body_GetPosition( b2BodyId bodyId );								 // :198
B2_API b2Rot b2Body_GetRotation( b2BodyId bodyId );									 // :201
B2_API b2Transform b2Body_GetTransform( b2BodyId bodyId );							 // :204
B2_API void b2Body_SetTransform( b2BodyId bodyId, b2Vec2 position, b2Rot rotation ); // :209
B2_API b2Vec2 b2Body_GetLocalPoint( b2BodyId bodyId, b2Vec2 worldPoint );			 // :212
B2_API b2Vec2 b2Body_GetWorldPoint( b2BodyId bodyId, b2Vec2 localPoint );			 // :215
B2_API b2Vec2 b2Body_GetLinearVelocity( b2BodyId bodyId );							 // :224
B2_API float b2Body_GetAngularVelocity( b2BodyId bodyId );							 // :227
B2_API void b2Body_SetLinearVelocity( b2BodyId bodyId, b2Vec2 linearVelocity );		 // :230
B2_API void b2Body_SetAngularVelocity( b2BodyId bodyId, float angularVelocity );	 // :233
B2_API void b2Body_ApplyForce( b2BodyId bodyId, b2Vec2 force, b2Vec2 point, bool wake );				  // :242
B2_API void b2Body_ApplyForceToCenter( b2BodyId bodyId, b2Vec2 force, bool wake );						  // :249
B2_API void b2Body_ApplyTorque( b2BodyId bodyId, float torque, bool wake );								  // :256
B2_API void b2Body_ApplyLinearImpulse( b2BodyId bodyId, b2Vec2 impulse, b2Vec2 point, bool wake );		  // :268
B2_API void b2Body_ApplyLinearImpulseToCenter( b2BodyId bodyId, b2Vec2 impulse, bool wake );			  // :277
B2_API void b2Body_ApplyAngularImpulse( b2BodyId bodyId, float impulse, bool wake );					  // :286
B2_API float b2Body_GetMass( b2BodyId bodyId );															  // :289
B2_API float b2Body_GetInertiaTensor( b2BodyId bodyId );												  // :292
B2_API b2Vec2 b2Body_GetLocalCenterOfMass( b2BodyId bodyId );											  // :295
B2_API b2Vec2 b2Body_GetWorldCenterOfMass( b2BodyId bodyId );											  // :298
B2_API b2MassData b2Body_GetMassData( b2BodyId bodyId );												  // :306
B2_API void b2Body_SetLinearDamping( b2BodyId bodyId, float linearDamping );							  // :323
B2_API float b2Body_GetLinearDamping( b2BodyId bodyId );												  // :326
B2_API void b2Body_SetAngularDamping( b2BodyId bodyId, float angularDamping );							  // :329
B2_API float b2Body_GetAngularDamping( b2BodyId bodyId );												  // :332
B2_API void b2Body_SetGravityScale( b2BodyId bodyId, float gravityScale );								  // :336
B2_API float b2Body_GetGravityScale( b2BodyId bodyId );													  // :339
B2_API bool b2Body_IsAwake( b2BodyId bodyId );															  // :342
B2_API bool b2Body_IsBullet( b2BodyId bodyId );															  // :381
B2_API void b2Body_EnableHitEvents( b2BodyId bodyId, bool enableHitEvents );								  // :385
B2_API int b2Body_GetShapeCount( b2BodyId bodyId );														  // :388
B2_API int b2Body_GetShapes( b2BodyId bodyId, b2ShapeId* shapeArray, int capacity );					  // :392
B2_API int b2Body_GetContactCapacity( b2BodyId bodyId );												  // :402
B2_API int b2Body_GetContactData( b2BodyId bodyId, b2ContactData* contactData, int capacity );			  // :405
B2_API b2AABB b2Body_ComputeAABB( b2BodyId bodyId );													  // :409

// ---- shapes (reference include/box2d/box2d.h:423-574) ----
B2_API b2ShapeId b2CreateCircleShape( b2BodyId bodyId, const b2ShapeDef* def, const b2Circle* circle );	   // :423
B2_API b2ShapeId b2CreateSegmentShape( b2BodyId bodyId, const b2ShapeDef* def, const b2Segment* segment );  // :428
B2_API b2ShapeId b2CreateCapsuleShape( b2BodyId bodyId, const b2ShapeDef* def, const b2Capsule* capsule );  // :433
B2_API b2ShapeId b2CreatePolygonShape( b2BodyId bodyId, const b2ShapeDef* def, const b2Polygon* polygon );  // :438
B2_API bool b2Shape_IsValid( b2ShapeId id );															   // :444
B2_API b2ShapeType b2Shape_GetType( b2ShapeId shapeId );												   // :447
B2_API b2BodyId b2Shape_GetBody( b2ShapeId shapeId );													   // :450
B2_API float b2Shape_GetDensity( b2ShapeId shapeId );													   // :468
B2_API void b2Shape_SetFriction( b2ShapeId shapeId, float friction );									   // :472
B2_API float b2Shape_GetFriction( b2ShapeId shapeId );													   // :475
B2_API void b2Shape_SetRestitution( b2ShapeId shapeId, float restitution );								   // :479
B2_API float b2Shape_GetRestitution( b2ShapeId shapeId );												   // :482
B2_API b2Filter b2Shape_GetFilter( b2ShapeId shapeId );													   // :485
B2_API void b2Shape_EnableContactEvents( b2ShapeId shapeId, bool flag );									   // :500
B2_API bool b2Shape_AreContactEventsEnabled( b2ShapeId shapeId );											   // :503
B2_API void b2Shape_EnableHitEvents( b2ShapeId shapeId, bool flag );										   // :515
B2_API bool b2Shape_AreHitEventsEnabled( b2ShapeId shapeId );												   // :518
B2_API b2Circle b2Shape_GetCircle( b2ShapeId shapeId );													   // :527
B2_API b2Segment b2Shape_GetSegment( b2ShapeId shapeId );												   // :530
B2_API b2Capsule b2Shape_GetCapsule( b2ShapeId shapeId );												   // :537
B2_API b2Polygon b2Shape_GetPolygon( b2ShapeId shapeId );												   // :540
B2_API int b2Shape_GetContactCapacity( b2ShapeId shapeId );												   // :565
B2_API int b2Shape_GetContactData( b2ShapeId shapeId, b2ContactData* contactData, int capacity );		   // :568
B2_API b2AABB b2Shape_GetAABB( b2ShapeId shapeId );														   // :571
