// Definition structs, enums and reporting structs of the public API.
//
// Field order and types follow the reference's include/box2d/types.h (task callbacks 30-48,
// b2WorldDef 56-109, b2BodyDef 128-192, b2Filter 197-213, b2ShapeType 227-235, b2ShapeDef
// 241-286, b2Profile 397-421, b2Counters 424-438, the joint definitions 446-857, b2ContactData
// 1010-1015) so that a definition filled in by unchanged application code means the same thing
// here.
#pragma once

#include "base.h"
#include "collision.h"
#include "id.h"
#include "math_functions.h"

#include <stdbool.h>
#include <stdint.h>

// Task plugin seam of the reference. Accepted and stored for source compatibility; the
// device pipeline does not call it (there is no host parallel-for on the hot path).
typedef void b2TaskCallback( int32_t startIndex, int32_t endIndex, uint32_t workerIndex, void* taskContext );
typedef void* b2EnqueueTaskCallback( b2TaskCallback* task, int32_t itemCount, int32_t minRange, void* taskContext,
									 void* userContext );
typedef void b2FinishTaskCallback( void* userTask, void* userContext );

typedef struct b2WorldDef
{
	b2Vec2 gravity;
	float restitutionThreshold;
	float contactPushoutVelocity;
	float hitEventThreshold;
	float contactHertz;
	float contactDampingRatio;
	float jointHertz;
	float jointDampingRatio;
	bool enableSleep;
	bool enableContinous;
	int32_t workerCount;
	b2EnqueueTaskCallback* enqueueTask;
	b2FinishTaskCallback* finishTask;
	void* userTaskContext;
	int32_t internalValue;
} b2WorldDef;

B2_API b2WorldDef b2DefaultWorldDef( void );

typedef enum b2BodyType
{
	b2_staticBody = 0,
	b2_kinematicBody = 1,
	b2_dynamicBody = 2,
	b2_bodyTypeCount,
} b2BodyType;

typedef struct b2BodyDef
{
	b2BodyType type;
	b2Vec2 position;
	b2Rot rotation;
	b2Vec2 linearVelocity;
	float angularVelocity;
	float linearDamping;
	float angularDamping;
	float gravityScale;
	float sleepThreshold;
	void* userData;
	bool enableSleep;
	bool isAwake;
	bool fixedRotation;
	bool isBullet;
	bool isEnabled;
	bool automaticMass;
	int32_t internalValue;
} b2BodyDef;

B2_API b2BodyDef b2DefaultBodyDef( void );

typedef struct b2Filter
{
	uint32_t categoryBits;
	uint32_t maskBits;
	int32_t groupIndex;
} b2Filter;

B2_API b2Filter b2DefaultFilter( void );

typedef enum b2ShapeType
{
	b2_circleShape,
	b2_capsuleShape,
	b2_segmentShape,
	b2_polygonShape,
	b2_smoothSegmentShape,
	b2_shapeTypeCount
} b2ShapeType;

typedef struct b2ShapeDef
{
	void* userData;
	float friction;
	float restitution;
	float density;
	b2Filter filter;
	uint32_t customColor;
	bool isSensor;
	bool enableSensorEvents;
	bool enableContactEvents;
	bool enableHitEvents;
	bool enablePreSolveEvents;
	bool forceContactCreation;
	int32_t internalValue;
} b2ShapeDef;

B2_API b2ShapeDef b2DefaultShapeDef( void );

// Per-stage wall-clock milliseconds of the last step.
typedef struct b2Profile
{
	float step;
	float pairs;
	float collide;
	float solve;
	float buildIslands;
	float solveConstraints;
	float prepareTasks;
	float solverTasks;
	float prepareConstraints;
	float integrateVelocities;
	float warmStart;
	float solveVelocities;
	float integratePositions;
	float relaxVelocities;
	float applyRestitution;
	float storeImpulses;
	float finalizeBodies;
	float splitIslands;
	float sleepIslands;
	float hitEvents;
	float broadphase;
	float continuous;
} b2Profile;

// Object counts; colorCounts[11] is the overflow set of the constraint graph.
typedef struct b2Counters
{
	int32_t staticBodyCount;
	int32_t bodyCount;
	int32_t shapeCount;
	int32_t contactCount;
	int32_t jointCount;
	int32_t islandCount;
	int32_t stackUsed;
	int32_t staticTreeHeight;
	int32_t treeHeight;
	int32_t byteCount;
	int32_t taskCount;
	int32_t colorCounts[12];
} b2Counters;

// Touching contact of a body or shape, as returned by b2Body_GetContactData.
typedef struct b2ContactData
{
	b2ShapeId shapeIdA;
	b2ShapeId shapeIdB;
	b2Manifold manifold;
} b2ContactData;

// ---- joints (reference include/box2d/types.h:446-857) ----

typedef enum b2JointType // :446
{
	b2_distanceJoint,
	b2_motorJoint,
	b2_mouseJoint,
	b2_prismaticJoint,
	b2_revoluteJoint,
	b2_weldJoint,
	b2_wheelJoint,
} b2JointType;

typedef struct b2DistanceJointDef // :464
{
	b2BodyId bodyIdA;
	b2BodyId bodyIdB;
	b2Vec2 localAnchorA;
	b2Vec2 localAnchorB;
	float length;
	bool enableSpring;
	float hertz;
	float dampingRatio;
	bool enableLimit;
	float minLength;
	float maxLength;
	bool enableMotor;
	float maxMotorForce;
	float motorSpeed;
	bool collideConnected;
	void* userData;
	int32_t internalValue;
} b2DistanceJointDef;

B2_API b2DistanceJointDef b2DefaultDistanceJointDef( void ); // :521

typedef struct b2MotorJointDef // :527
{
	b2BodyId bodyIdA;
	b2BodyId bodyIdB;
	b2Vec2 linearOffset;
	float angularOffset;
	float maxForce;
	float maxTorque;
	float correctionFactor;
	bool collideConnected;
	void* userData;
	int32_t internalValue;
} b2MotorJointDef;

B2_API b2MotorJointDef b2DefaultMotorJointDef( void ); // :562

typedef struct b2MouseJointDef // :569
{
	b2BodyId bodyIdA;
	b2BodyId bodyIdB;
	b2Vec2 target;
	float hertz;
	float dampingRatio;
	float maxForce;
	bool collideConnected;
	void* userData;
	int32_t internalValue;
} b2MouseJointDef;

B2_API b2MouseJointDef b2DefaultMouseJointDef( void ); // :601

typedef struct b2PrismaticJointDef // :610
{
	b2BodyId bodyIdA;
	b2BodyId bodyIdB;
	b2Vec2 localAnchorA;
	b2Vec2 localAnchorB;
	b2Vec2 localAxisA;
	float referenceAngle;
	bool enableSpring;
	float hertz;
	float dampingRatio;
	bool enableLimit;
	float lowerTranslation;
	float upperTranslation;
	bool enableMotor;
	float maxMotorForce;
	float motorSpeed;
	bool collideConnected;
	void* userData;
	int32_t internalValue;
} b2PrismaticJointDef;

B2_API b2PrismaticJointDef b2DefaultPrismaticJointDef( void ); // :669

typedef struct b2RevoluteJointDef // :683
{
	b2BodyId bodyIdA;
	b2BodyId bodyIdB;
	b2Vec2 localAnchorA;
	b2Vec2 localAnchorB;
	float referenceAngle;
	bool enableSpring;
	float hertz;
	float dampingRatio;
	bool enableLimit;
	float lowerAngle;
	float upperAngle;
	bool enableMotor;
	float maxMotorTorque;
	float motorSpeed;
	float drawSize;
	bool collideConnected;
	void* userData;
	int32_t internalValue;
} b2RevoluteJointDef;

B2_API b2RevoluteJointDef b2DefaultRevoluteJointDef( void ); // :743

typedef struct b2WeldJointDef // :751
{
	b2BodyId bodyIdA;
	b2BodyId bodyIdB;
	b2Vec2 localAnchorA;
	b2Vec2 localAnchorB;
	float referenceAngle;
	float linearHertz;
	float angularHertz;
	float linearDampingRatio;
	float angularDampingRatio;
	bool collideConnected;
	void* userData;
	int32_t internalValue;
} b2WeldJointDef;

B2_API b2WeldJointDef b2DefaultWeldJointDef( void ); // :792

typedef struct b2WheelJointDef // :801
{
	b2BodyId bodyIdA;
	b2BodyId bodyIdB;
	b2Vec2 localAnchorA;
	b2Vec2 localAnchorB;
	b2Vec2 localAxisA;
	bool enableSpring;
	float hertz;
	float dampingRatio;
	bool enableLimit;
	float lowerTranslation;
	float upperTranslation;
	bool enableMotor;
	float maxMotorTorque;
	float motorSpeed;
	bool collideConnected;
	void* userData;
	int32_t internalValue;
} b2WheelJointDef;

B2_API b2WheelJointDef b2DefaultWheelJointDef( void ); // :857

// ---- events of the last step (reference include/box2d/types.h:876-1005) ----
// Filled by b2World_Step and valid until the next step; the arrays are owned by the world.

// a shape began (end: stopped) overlapping a sensor shape
typedef struct b2SensorBeginTouchEvent
{
	b2ShapeId sensorShapeId;
	b2ShapeId visitorShapeId;
} b2SensorBeginTouchEvent;

typedef struct b2SensorEndTouchEvent
{
	b2ShapeId sensorShapeId;
	b2ShapeId visitorShapeId;
} b2SensorEndTouchEvent;

typedef struct b2SensorEvents
{
	b2SensorBeginTouchEvent* beginEvents;
	b2SensorEndTouchEvent* endEvents;
	int32_t beginCount;
	int32_t endCount;
} b2SensorEvents;

// two shapes began (end: stopped) touching
typedef struct b2ContactBeginTouchEvent
{
	b2ShapeId shapeIdA;
	b2ShapeId shapeIdB;
} b2ContactBeginTouchEvent;

typedef struct b2ContactEndTouchEvent
{
	b2ShapeId shapeIdA;
	b2ShapeId shapeIdB;
} b2ContactEndTouchEvent;

// two shapes met faster than the world's hit event threshold: the point of the fastest
// approach, the manifold normal (A to B) and that approach speed (positive)
typedef struct b2ContactHitEvent
{
	b2ShapeId shapeIdA;
	b2ShapeId shapeIdB;
	b2Vec2 point;
	b2Vec2 normal;
	float approachSpeed;
} b2ContactHitEvent;

typedef struct b2ContactEvents
{
	b2ContactBeginTouchEvent* beginEvents;
	b2ContactEndTouchEvent* endEvents;
	b2ContactHitEvent* hitEvents;
	int32_t beginCount;
	int32_t endCount;
	int32_t hitCount;
} b2ContactEvents;

// one per body the step moved (every awake body, in awake order): its new transform
typedef struct b2BodyMoveEvent
{
	b2Transform transform;
	b2BodyId bodyId;
	void* userData;
	bool fellAsleep;
} b2BodyMoveEvent;

typedef struct b2BodyEvents
{
	b2BodyMoveEvent* moveEvents;
	int32_t moveCount;
} b2BodyEvents;

// User filter and pre-solve hooks. They would run inside the data-parallel narrowphase;
// this library refuses them (see b2World_SetCustomFilterCallback).
typedef bool b2CustomFilterFcn( b2ShapeId shapeIdA, b2ShapeId shapeIdB, void* context );
typedef bool b2PreSolveFcn( b2ShapeId shapeIdA, b2ShapeId shapeIdB, b2Manifold* manifold, void* context );
