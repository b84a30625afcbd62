// CPU ORACLE — test infrastructure only.
//
// A plain-C restatement of the reference's b2World_Step pipeline, used by tests/, by
// __graft_entry__.smoke() and by bench.py's cpu_baseline leg as the checker of the HIP
// path. It is never linked into, loaded by or called from the product (box2c_amd).
//
// It implements the same public C API (include/box2d/box2d.h) so one scene file drives the
// reference, this oracle and the product alike. Single-threaded; every function cites the
// reference file:line it restates. Contracts that it reproduces exactly:
//  - id allocation (LIFO pools, src/id_pool.c:21-47), swap-remove arrays
//    (src/block_array.c:244-257), head-inserted shape and contact-edge lists;
//  - dynamic-tree structure (insert/rotate/enlarge/rebuild) and DFS query order;
//  - move-array / pair-list / contact-creation order, greedy graph colouring;
//  - float arithmetic of the narrowphase and of every solver stage, op for op, including
//    the differences between the 8-wide graph-colour path and the scalar overflow path.
//
// When built with ORACLE_SOLVER_F64 the solver stages, integrators and finalize run in
// double (the "FP64 build of the same step"): manifolds and the contact set still come from
// the float narrowphase; body state is stored in double between steps.
#pragma once

#include "box2d/box2d.h"
#include "b2mi.h"

#include <stdbool.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#if ORACLE_SOLVER_F64
typedef double real;
#else
typedef float real;
#endif

#define OR_NULL ( -1 )

// constants of the reference (src/core.h:110-151)
#define OR_LINEAR_SLOP 0.005f
#define OR_SPECULATIVE ( 4.0f * OR_LINEAR_SLOP )
#define OR_AABB_MARGIN 0.1f
#define OR_HUGE ( 100000.0f )
#define OR_MAX_TRANSLATION 4.0f
#define OR_MAX_ROTATION ( 0.25f * b2_pi )
#define OR_GRAPH_COLORS 12
#define OR_OVERFLOW 11
#define OR_MAX_WORLDS 128
#define OR_COOKIE 1152023

// ---- growable arrays ----
#define OR_ARRAY( T ) \
	struct \
	{ \
		T* data; \
		int count; \
		int capacity; \
	}
#define OR_RESERVE( a, n ) \
	do \
	{ \
		if ( ( n ) > ( a ).capacity ) \
		{ \
			int cap_ = ( a ).capacity ? ( a ).capacity : 16; \
			while ( cap_ < ( n ) ) \
				cap_ *= 2; \
			( a ).data = realloc( ( a ).data, (size_t)cap_ * sizeof( *( a ).data ) ); \
			( a ).capacity = cap_; \
		} \
	} while ( 0 )
#define OR_PUSH( a, v ) \
	do \
	{ \
		OR_RESERVE( a, ( a ).count + 1 ); \
		( a ).data[( a ).count++] = ( v ); \
	} while ( 0 )
#define OR_FREE( a ) \
	do \
	{ \
		free( ( a ).data ); \
		( a ).data = NULL; \
		( a ).count = ( a ).capacity = 0; \
	} while ( 0 )

typedef OR_ARRAY( int ) orIntArray;

// LIFO id pool; freed ids are always pushed (src/id_pool.c:36-47: the "return to
// nextIndex" branch can never fire because ids are < nextIndex).
typedef struct orIdPool
{
	orIntArray freeArray;
	int nextIndex;
} orIdPool;

int or_alloc_id( orIdPool* pool );
void or_free_id( orIdPool* pool, int id );

// ---- bitset (src/bitset.h) ----
typedef struct orBitSet
{
	uint64_t* bits;
	uint32_t blockCount;
	uint32_t blockCapacity;
} orBitSet;

void or_bitset_set_count_and_clear( orBitSet* s, uint32_t bitCount );
void or_bitset_grow( orBitSet* s, uint32_t blockCount );
static inline void or_bitset_set( orBitSet* s, uint32_t i )
{
	uint32_t b = i / 64;
	if ( b >= s->blockCount )
		or_bitset_grow( s, b + 1 );
	s->bits[b] |= (uint64_t)1 << ( i % 64 );
}
static inline void or_bitset_clear( orBitSet* s, uint32_t i )
{
	uint32_t b = i / 64;
	if ( b < s->blockCount )
		s->bits[b] &= ~( (uint64_t)1 << ( i % 64 ) );
}
static inline bool or_bitset_get( const orBitSet* s, uint32_t i )
{
	uint32_t b = i / 64;
	if ( b >= s->blockCount )
		return false;
	return ( s->bits[b] >> ( i % 64 ) ) & 1;
}
void or_bitset_free( orBitSet* s );

// ---- 64-bit key hash set (membership only; src/table.c) ----
typedef struct orKeySet
{
	uint64_t* keys; // 0 = empty slot
	uint32_t capacity;
	uint32_t count;
} orKeySet;

bool or_keyset_add( orKeySet* s, uint64_t key ); // returns true if already present
bool or_keyset_remove( orKeySet* s, uint64_t key );
bool or_keyset_contains( const orKeySet* s, uint64_t key );
void or_keyset_clear( orKeySet* s );
void or_keyset_free( orKeySet* s );

static inline uint64_t or_pair_key( int a, int b )
{
	return a < b ? ( (uint64_t)a << 32 | (uint64_t)b ) : ( (uint64_t)b << 32 | (uint64_t)a );
}

// ---- dynamic tree (src/dynamic_tree.c, node include/box2d/collision.h:600-635) ----
typedef struct orTreeNode
{
	b2AABB aabb;
	uint32_t categoryBits;
	int32_t parent; // also the free-list link
	int32_t child1;
	int32_t child2;
	int32_t userData;
	int16_t height; // -1 free, 0 leaf
	bool enlarged;
} orTreeNode;

typedef struct orTree
{
	orTreeNode* nodes;
	int root;
	int nodeCount;
	int nodeCapacity;
	int freeList;
	int proxyCount;
	int* leafIndices;
	b2Vec2* leafCenters;
	int rebuildCapacity;
} orTree;

void or_tree_create( orTree* tree );
void or_tree_destroy( orTree* tree );
int or_tree_create_proxy( orTree* tree, b2AABB aabb, uint32_t categoryBits, int userData );
void or_tree_destroy_proxy( orTree* tree, int proxyId );
void or_tree_move_proxy( orTree* tree, int proxyId, b2AABB aabb );
void or_tree_enlarge_proxy( orTree* tree, int proxyId, b2AABB aabb );
int or_tree_rebuild( orTree* tree, bool fullBuild );
int or_tree_height( const orTree* tree );
typedef bool orTreeQueryFcn( int proxyId, int userData, void* context );
void or_tree_query( const orTree* tree, b2AABB aabb, uint32_t maskBits, orTreeQueryFcn* fcn, void* context );

static inline bool or_aabb_overlaps( b2AABB a, b2AABB b )
{
	// src/aabb.h:59-71
	float d1x = b.lowerBound.x - a.upperBound.x, d1y = b.lowerBound.y - a.upperBound.y;
	float d2x = a.lowerBound.x - b.upperBound.x, d2y = a.lowerBound.y - b.upperBound.y;
	if ( d1x > 0.0f || d1y > 0.0f )
		return false;
	if ( d2x > 0.0f || d2y > 0.0f )
		return false;
	return true;
}

// ---- manifolds ----
b2Manifold or_collide_shapes( int typeA, const void* geomA, b2Transform xfA, int typeB, const void* geomB, b2Transform xfB );
bool or_has_manifold_fcn( int typeA, int typeB, bool* primary );

// ---- world model ----
typedef struct orBody
{
	void* userData;
	int setIndex; // OR_NULL when free; 0 static set, 2 awake set
	int localIndex;
	int headContactKey;
	int contactCount;
	int headShapeId;
	int shapeCount;
	int headJointKey;
	int jointCount;
	float sleepThreshold;
	int id;
	b2BodyType type;
	uint16_t revision;
	bool enableSleep;
	bool fixedRotation;
	bool isSpeedCapped;
	bool automaticMass;
} orBody;

typedef struct orVec2r
{
	real x, y;
} orVec2r;

typedef struct orBodySim
{
	b2Transform transform;
	b2Vec2 center;
	b2Rot rotation0;
	b2Vec2 center0;
	b2Vec2 localCenter;
	b2Vec2 force;
	float torque;
	float mass, invMass;
	float inertia, invInertia;
	float minExtent, maxExtent;
	float linearDamping, angularDamping, gravityScale;
	int bodyId;
	bool isFast, isBullet, isSpeedCapped, enlargeAABB;
#if ORACLE_SOLVER_F64
	// double-precision shadow of the pose carried across steps
	orVec2r centerR;
	orVec2r qR; // (c, s)
#endif
} orBodySim;

// src/body.h:76-88 (flags kept for layout parity with the 8-wide gather)
typedef struct orBodyState
{
	orVec2r v;
	real w;
	int flags;
	orVec2r dp;
	orVec2r dq; // (c, s)
} orBodyState;

typedef struct orShape
{
	int id; // OR_NULL when free
	int bodyId;
	int prevShapeId;
	int nextShapeId;
	b2ShapeType type;
	float density;
	float friction;
	float restitution;
	b2AABB aabb;
	b2AABB fatAABB;
	b2Vec2 localCentroid;
	int proxyKey;
	b2Filter filter;
	void* userData;
	union
	{
		b2Capsule capsule;
		b2Circle circle;
		b2Polygon polygon;
		b2Segment segment;
	};
	uint16_t revision;
	bool isSensor;
	bool enableContactEvents;
	bool enableHitEvents;
	bool enlargedAABB;
	bool isFast;
} orShape;

typedef struct orContactEdge
{
	int bodyId;
	int prevKey;
	int nextKey;
} orContactEdge;

enum
{
	OR_CONTACT_TOUCHING = 0x1,
	OR_CONTACT_ENABLE_CONTACT_EVENTS = 0x20,
};

// src/contact.h:96-115
enum
{
	OR_SIM_TOUCHING = 0x00010000,
	OR_SIM_DISJOINT = 0x00020000,
	OR_SIM_STARTED = 0x00040000,
	OR_SIM_STOPPED = 0x00080000,
	OR_SIM_HIT_EVENT = 0x00100000,
};

typedef struct orContact
{
	int setIndex; // OR_NULL when free
	int colorIndex;
	int localIndex;
	orContactEdge edges[2];
	int shapeIdA, shapeIdB;
	int contactId;
	uint32_t flags;
} orContact;

typedef struct orContactSim
{
	int contactId;
	int bodySimIndexA, bodySimIndexB;
	int shapeIdA, shapeIdB;
	float invMassA, invIA, invMassB, invIB;
	b2Manifold manifold;
	float friction, restitution;
	uint32_t simFlags;
} orContactSim;

typedef OR_ARRAY( orContactSim ) orContactSimArray;

typedef struct orSoftness
{
	real biasRate, massScale, impulseScale;
} orSoftness;

// ---- joints (src/joint.h:13-178, 182-226) ----
typedef struct orMat22r
{
	orVec2r cx, cy;
} orMat22r;

typedef struct orJointEdge
{
	int bodyId;
	int prevKey;
	int nextKey;
} orJointEdge;

typedef struct orJoint
{
	void* userData;
	int setIndex; // OR_NULL when free; 0 static set, 2 awake set
	int colorIndex;
	int localIndex;
	orJointEdge edges[2];
	int jointId;
	b2JointType type;
	uint16_t revision;
	float drawSize;
	bool collideConnected;
} orJoint;

typedef struct orJointSim
{
	int jointId;
	int bodyIdA, bodyIdB;
	b2JointType type;
	b2Vec2 localOriginAnchorA, localOriginAnchorB;
	real invMassA, invMassB, invIA, invIB;
	int indexA, indexB; // awake sim index, OR_NULL for a static body
	bool enableSpring, enableLimit, enableMotor;
	// parameters
	float referenceAngle; // revolute, prismatic, weld
	float hertz, dampingRatio; // spring of revolute, prismatic, wheel, distance; mouse
	float linearHertz, linearDampingRatio, angularHertz, angularDampingRatio; // weld
	float lower, upper; // revolute angles, prismatic and wheel translations, distance lengths
	float maxMotor;		// max motor torque (revolute, wheel) or force (prismatic, distance)
	float motorSpeed;
	float length; // distance
	b2Vec2 localAxisA; // prismatic, wheel
	b2Vec2 linearOffset; // motor
	float angularOffset, maxForce, maxTorque, correctionFactor; // motor, mouse (maxForce)
	b2Vec2 target; // mouse
	// accumulated impulses
	orVec2r linearImpulse; // revolute, weld, motor, mouse; prismatic's 2D impulse
	real angularImpulse; // weld, motor, mouse
	real impulse; // distance axial, wheel perpendicular
	real springImpulse, motorImpulse, lowerImpulse, upperImpulse;
	// per-step terms of prepare
	orVec2r anchorA, anchorB, deltaCenter, axisA;
	real deltaAngle, axialMass, perpMass, motorMass, angularMass;
	orMat22r linearMass;
	orSoftness springSoftness, linearSoftness, angularSoftness;
} orJointSim;

typedef OR_ARRAY( orJointSim ) orJointSimArray;
typedef OR_ARRAY( orBodySim ) orBodySimArray;
typedef OR_ARRAY( orBodyState ) orBodyStateArray;

typedef struct orGraphColor
{
	orBitSet bodySet;
	orContactSimArray contacts;
	orJointSimArray joints;
} orGraphColor;

typedef struct orWorld
{
	int worldId; // OR_NULL when free
	uint16_t revision;
	bool inUse;
	bool locked;

	OR_ARRAY( orBody ) bodies;
	orIdPool bodyIdPool;
	orBodySimArray staticSims;
	orBodySimArray awakeSims;
	orBodyStateArray awakeStates;
	orContactSimArray awakeContacts; // non-touching

	OR_ARRAY( orShape ) shapes;
	orIdPool shapeIdPool;

	OR_ARRAY( orContact ) contacts;
	orIdPool contactIdPool;

	OR_ARRAY( orJoint ) joints;
	orIdPool jointIdPool;
	orJointSimArray staticJoints; // joints between two static bodies (the static set's)
	float inv_h;				  // of the last step (src/world.c:731)

	orGraphColor colors[OR_GRAPH_COLORS];

	// broadphase (src/broad_phase.h:26-53)
	orTree trees[3];
	orBitSet moveSet; // indexed by proxyKey
	orIntArray moveArray;
	orKeySet pairSet;

	b2Vec2 gravity;
	float hitEventThreshold, restitutionThreshold, contactPushoutVelocity;
	float contactHertz, contactDampingRatio, jointHertz, jointDampingRatio;
	bool enableSleep, enableWarmStarting, enableContinuous;
	int stepIndex;
	int toiHits; // time-of-impact events so far (a test probe)

	b2Profile profile;

	// events of the last step (src/world.h: bodyMoveEventArray, sensor/contact event arrays)
	OR_ARRAY( b2BodyMoveEvent ) moveEvents;
	OR_ARRAY( b2SensorBeginTouchEvent ) sensorBegin;
	OR_ARRAY( b2SensorEndTouchEvent ) sensorEnd;
	OR_ARRAY( b2ContactBeginTouchEvent ) contactBegin;
	OR_ARRAY( b2ContactEndTouchEvent ) contactEnd;
	OR_ARRAY( b2ContactHitEvent ) contactHit;
} orWorld;

typedef struct orStepContext
{
	float dt, inv_dt, h, inv_h;
	int subStepCount;
	orSoftness contactSoftness, staticSoftness, jointSoftness;
	float restitutionThreshold;
	bool enableWarmStarting;
} orStepContext;

extern orWorld or_worlds[OR_MAX_WORLDS];

orWorld* or_get_world( b2WorldId id );
orWorld* or_get_world_index( int index );
orBody* or_get_body_full( orWorld* world, b2BodyId id );
orBodySim* or_get_body_sim( orWorld* world, orBody* body );
orShape* or_get_shape_full( orWorld* world, b2ShapeId id );

b2AABB or_compute_shape_aabb( const orShape* shape, b2Transform xf );
void or_update_body_mass( orWorld* world, orBody* body );

void or_update_pairs( orWorld* world );
void or_create_contact( orWorld* world, orShape* shapeA, orShape* shapeB );
void or_destroy_contact( orWorld* world, orContact* contact );
void or_add_contact_to_graph( orWorld* world, orContactSim* contactSim, orContact* contact );
void or_remove_contact_from_graph( orWorld* world, int bodyIdA, int bodyIdB, int colorIndex, int localIndex );
void or_collide( orWorld* world, orStepContext* ctx );
bool or_should_bodies_collide( orWorld* world, orBody* bodyA, orBody* bodyB );
void or_destroy_joint_internal( orWorld* world, orJoint* joint );
void or_prepare_joint( orWorld* world, orJointSim* joint, const orStepContext* ctx );
void or_warm_start_joint( orJointSim* joint, orBodyState* states );
void or_solve_joint( orJointSim* joint, orBodyState* states, const orStepContext* ctx, bool useBias );
orSoftness or_make_softness( real hertz, real zeta, real h );
void or_solve( orWorld* world, orStepContext* ctx );
void or_assert_fail( const char* what, const char* file, int line );

#define OR_ASSERT( c ) \
	do \
	{ \
		if ( !( c ) ) \
			or_assert_fail( #c, __FILE__, __LINE__ ); \
	} while ( 0 )
