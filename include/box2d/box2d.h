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
B2_API int b2Body_GetJointCount( b2BodyId bodyId );										  // :395
B2_API int b2Body_GetJoints( b2BodyId bodyId, b2JointId* jointArray, int capacity );					  // :399

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

// ---- joints (reference include/box2d/box2d.h:599-1048) ----
// Joints are solved by the step kernels beside the contacts, colour by colour, with the overflow
// joints before the overflow contacts; a joint needs at least one dynamic body.
B2_API void b2DestroyJoint( b2JointId jointId );										  // :605
B2_API bool b2Joint_IsValid( b2JointId id );											  // :608
B2_API b2JointType b2Joint_GetType( b2JointId jointId );								  // :611
B2_API b2BodyId b2Joint_GetBodyA( b2JointId jointId );									  // :614
B2_API b2BodyId b2Joint_GetBodyB( b2JointId jointId );									  // :617
B2_API b2Vec2 b2Joint_GetLocalAnchorA( b2JointId jointId );							  // :620
B2_API b2Vec2 b2Joint_GetLocalAnchorB( b2JointId jointId );							  // :623
B2_API void b2Joint_SetCollideConnected( b2JointId jointId, bool shouldCollide );		  // :626
B2_API bool b2Joint_GetCollideConnected( b2JointId jointId );							  // :629
B2_API void b2Joint_SetUserData( b2JointId jointId, void* userData );					  // :632
B2_API void* b2Joint_GetUserData( b2JointId jointId );									  // :635
B2_API void b2Joint_WakeBodies( b2JointId jointId );									  // :638
B2_API b2Vec2 b2Joint_GetConstraintForce( b2JointId jointId );							  // :641
B2_API float b2Joint_GetConstraintTorque( b2JointId jointId );							  // :644

B2_API b2JointId b2CreateDistanceJoint( b2WorldId worldId, const b2DistanceJointDef* def );	  // :654
B2_API b2JointId b2CreateMotorJoint( b2WorldId worldId, const b2MotorJointDef* def );		  // :736
B2_API b2JointId b2CreateMouseJoint( b2WorldId worldId, const b2MouseJointDef* def );		  // :781
B2_API b2JointId b2CreatePrismaticJoint( b2WorldId worldId, const b2PrismaticJointDef* def ); // :820
B2_API b2JointId b2CreateRevoluteJoint( b2WorldId worldId, const b2RevoluteJointDef* def );	  // :891
B2_API b2JointId b2CreateWeldJoint( b2WorldId worldId, const b2WeldJointDef* def );		  // :963
B2_API b2JointId b2CreateWheelJoint( b2WorldId worldId, const b2WheelJointDef* def );		  // :1003

// revolute joint (reference include/box2d/box2d.h:881-946)
B2_API void b2RevoluteJoint_EnableSpring( b2JointId jointId, bool enableSpring );		// :894
B2_API void b2RevoluteJoint_SetSpringHertz( b2JointId jointId, float hertz );			// :897
B2_API float b2RevoluteJoint_GetSpringHertz( b2JointId jointId );						// :900
B2_API void b2RevoluteJoint_SetSpringDampingRatio( b2JointId jointId, float dampingRatio ); // :903
B2_API float b2RevoluteJoint_GetSpringDampingRatio( b2JointId jointId );				// :906
B2_API float b2RevoluteJoint_GetAngle( b2JointId jointId );								// :910
B2_API void b2RevoluteJoint_EnableLimit( b2JointId jointId, bool enableLimit );		// :913
B2_API bool b2RevoluteJoint_IsLimitEnabled( b2JointId jointId );						// :916
B2_API float b2RevoluteJoint_GetLowerLimit( b2JointId jointId );						// :919
B2_API float b2RevoluteJoint_GetUpperLimit( b2JointId jointId );						// :922
B2_API void b2RevoluteJoint_SetLimits( b2JointId jointId, float lower, float upper );	// :925
B2_API void b2RevoluteJoint_EnableMotor( b2JointId jointId, bool enableMotor );		// :928
B2_API bool b2RevoluteJoint_IsMotorEnabled( b2JointId jointId );						// :931
B2_API void b2RevoluteJoint_SetMotorSpeed( b2JointId jointId, float motorSpeed );		// :934
B2_API float b2RevoluteJoint_GetMotorSpeed( b2JointId jointId );						// :937
B2_API float b2RevoluteJoint_GetMotorTorque( b2JointId jointId );						// :940
B2_API void b2RevoluteJoint_SetMaxMotorTorque( b2JointId jointId, float torque );		// :943
B2_API float b2RevoluteJoint_GetMaxMotorTorque( b2JointId jointId );					// :946

// weld joint (reference include/box2d/box2d.h:951-987)
B2_API void b2WeldJoint_SetLinearHertz( b2JointId jointId, float hertz );				// :966
B2_API float b2WeldJoint_GetLinearHertz( b2JointId jointId );							// :969
B2_API void b2WeldJoint_SetLinearDampingRatio( b2JointId jointId, float dampingRatio ); // :972
B2_API float b2WeldJoint_GetLinearDampingRatio( b2JointId jointId );					// :975
B2_API void b2WeldJoint_SetAngularHertz( b2JointId jointId, float hertz );				// :978
B2_API float b2WeldJoint_GetAngularHertz( b2JointId jointId );							// :981
B2_API void b2WeldJoint_SetAngularDampingRatio( b2JointId jointId, float dampingRatio ); // :984
B2_API float b2WeldJoint_GetAngularDampingRatio( b2JointId jointId );					// :987

// distance joint (reference include/box2d/box2d.h:649-720)
B2_API void b2DistanceJoint_SetLength( b2JointId jointId, float length ); // :659
B2_API float b2DistanceJoint_GetLength( b2JointId jointId ); // :662
B2_API void b2DistanceJoint_EnableSpring( b2JointId jointId, bool enableSpring ); // :665
B2_API bool b2DistanceJoint_IsSpringEnabled( b2JointId jointId ); // :668
B2_API void b2DistanceJoint_SetSpringHertz( b2JointId jointId, float hertz ); // :671
B2_API void b2DistanceJoint_SetSpringDampingRatio( b2JointId jointId, float dampingRatio ); // :674
B2_API float b2DistanceJoint_GetHertz( b2JointId jointId ); // :677
B2_API float b2DistanceJoint_GetDampingRatio( b2JointId jointId ); // :680
B2_API void b2DistanceJoint_EnableLimit( b2JointId jointId, bool enableLimit ); // :684
B2_API bool b2DistanceJoint_IsLimitEnabled( b2JointId jointId ); // :687
B2_API void b2DistanceJoint_SetLengthRange( b2JointId jointId, float minLength, float maxLength ); // :690
B2_API float b2DistanceJoint_GetMinLength( b2JointId jointId ); // :693
B2_API float b2DistanceJoint_GetMaxLength( b2JointId jointId ); // :696
B2_API float b2DistanceJoint_GetCurrentLength( b2JointId jointId ); // :699
B2_API void b2DistanceJoint_EnableMotor( b2JointId jointId, bool enableMotor ); // :702
B2_API bool b2DistanceJoint_IsMotorEnabled( b2JointId jointId ); // :705
B2_API void b2DistanceJoint_SetMotorSpeed( b2JointId jointId, float motorSpeed ); // :708
B2_API float b2DistanceJoint_GetMotorSpeed( b2JointId jointId ); // :711
B2_API void b2DistanceJoint_SetMaxMotorForce( b2JointId jointId, float force ); // :714
B2_API float b2DistanceJoint_GetMaxMotorForce( b2JointId jointId ); // :717
B2_API float b2DistanceJoint_GetMotorForce( b2JointId jointId ); // :720

// motor joint (reference include/box2d/box2d.h:725-766)
B2_API void b2MotorJoint_SetLinearOffset( b2JointId jointId, b2Vec2 linearOffset ); // :739
B2_API b2Vec2 b2MotorJoint_GetLinearOffset( b2JointId jointId ); // :742
B2_API void b2MotorJoint_SetAngularOffset( b2JointId jointId, float angularOffset ); // :745
B2_API float b2MotorJoint_GetAngularOffset( b2JointId jointId ); // :748
B2_API void b2MotorJoint_SetMaxForce( b2JointId jointId, float maxForce ); // :751
B2_API float b2MotorJoint_GetMaxForce( b2JointId jointId ); // :754
B2_API void b2MotorJoint_SetMaxTorque( b2JointId jointId, float maxTorque ); // :757
B2_API float b2MotorJoint_GetMaxTorque( b2JointId jointId ); // :760
B2_API void b2MotorJoint_SetCorrectionFactor( b2JointId jointId, float correctionFactor ); // :763
B2_API float b2MotorJoint_GetCorrectionFactor( b2JointId jointId ); // :766

// mouse joint (reference include/box2d/box2d.h:771-805)
B2_API void b2MouseJoint_SetTarget( b2JointId jointId, b2Vec2 target ); // :784
B2_API b2Vec2 b2MouseJoint_GetTarget( b2JointId jointId ); // :787
B2_API void b2MouseJoint_SetSpringHertz( b2JointId jointId, float hertz ); // :790
B2_API float b2MouseJoint_GetSpringHertz( b2JointId jointId ); // :793
B2_API void b2MouseJoint_SetSpringDampingRatio( b2JointId jointId, float dampingRatio ); // :796
B2_API float b2MouseJoint_GetSpringDampingRatio( b2JointId jointId ); // :799
B2_API void b2MouseJoint_SetMaxForce( b2JointId jointId, float maxForce ); // :802
B2_API float b2MouseJoint_GetMaxForce( b2JointId jointId ); // :805

// prismatic joint (reference include/box2d/box2d.h:810-876)
B2_API void b2PrismaticJoint_EnableSpring( b2JointId jointId, bool enableSpring ); // :823
B2_API bool b2PrismaticJoint_IsSpringEnabled( b2JointId jointId ); // :826
B2_API void b2PrismaticJoint_SetSpringHertz( b2JointId jointId, float hertz ); // :831
B2_API float b2PrismaticJoint_GetSpringHertz( b2JointId jointId ); // :834
B2_API void b2PrismaticJoint_SetSpringDampingRatio( b2JointId jointId, float dampingRatio ); // :837
B2_API float b2PrismaticJoint_GetSpringDampingRatio( b2JointId jointId ); // :840
B2_API void b2PrismaticJoint_EnableLimit( b2JointId jointId, bool enableLimit ); // :843
B2_API bool b2PrismaticJoint_IsLimitEnabled( b2JointId jointId ); // :846
B2_API float b2PrismaticJoint_GetLowerLimit( b2JointId jointId ); // :849
B2_API float b2PrismaticJoint_GetUpperLimit( b2JointId jointId ); // :852
B2_API void b2PrismaticJoint_SetLimits( b2JointId jointId, float lower, float upper ); // :855
B2_API void b2PrismaticJoint_EnableMotor( b2JointId jointId, bool enableMotor ); // :858
B2_API bool b2PrismaticJoint_IsMotorEnabled( b2JointId jointId ); // :861
B2_API void b2PrismaticJoint_SetMotorSpeed( b2JointId jointId, float motorSpeed ); // :864
B2_API float b2PrismaticJoint_GetMotorSpeed( b2JointId jointId ); // :867
B2_API void b2PrismaticJoint_SetMaxMotorForce( b2JointId jointId, float force ); // :870
B2_API float b2PrismaticJoint_GetMaxMotorForce( b2JointId jointId ); // :873
B2_API float b2PrismaticJoint_GetMotorForce( b2JointId jointId ); // :876

// wheel joint (reference include/box2d/box2d.h:992-1048)
B2_API void b2WheelJoint_EnableSpring( b2JointId jointId, bool enableSpring ); // :1006
B2_API bool b2WheelJoint_IsSpringEnabled( b2JointId jointId ); // :1009
B2_API void b2WheelJoint_SetSpringHertz( b2JointId jointId, float hertz ); // :1012
B2_API float b2WheelJoint_GetSpringHertz( b2JointId jointId ); // :1015
B2_API void b2WheelJoint_SetSpringDampingRatio( b2JointId jointId, float dampingRatio ); // :1018
B2_API float b2WheelJoint_GetSpringDampingRatio( b2JointId jointId ); // :1021
B2_API void b2WheelJoint_EnableLimit( b2JointId jointId, bool enableLimit ); // :1024
B2_API bool b2WheelJoint_IsLimitEnabled( b2JointId jointId ); // :1027
B2_API float b2WheelJoint_GetLowerLimit( b2JointId jointId ); // :1030
B2_API float b2WheelJoint_GetUpperLimit( b2JointId jointId ); // :1033
B2_API void b2WheelJoint_SetLimits( b2JointId jointId, float lower, float upper ); // :1036
B2_API void b2WheelJoint_EnableMotor( b2JointId jointId, bool enableMotor ); // :1039
B2_API bool b2WheelJoint_IsMotorEnabled( b2JointId jointId ); // :1042
B2_API void b2WheelJoint_SetMotorSpeed( b2JointId jointId, float motorSpeed ); // :1045
B2_API float b2WheelJoint_GetMotorSpeed( b2JointId jointId ); // :1048
B2_API void b2WheelJoint_SetMaxMotorTorque( b2JointId jointId, float torque ); // :1051
B2_API float b2WheelJoint_GetMaxMotorTorque( b2JointId jointId ); // :1054
B2_API float b2WheelJoint_GetMotorTorque( b2JointId jointId ); // :1057
