// Device-resident data layout of the step pipeline, shared by the host driver (world.cpp)
// and the HIP kernels (kernels.hip). Plain-old-data only.
//
// One DevWorld per world holds pointers to that world's device arrays; a step launches one
// narrowphase kernel over (world, contactId) pairs and one solver kernel over work items.
// A work item is a union of whole islands (connected components of non-static bodies through
// touching contacts), so no two work items write the same body and a workgroup can run every
// stage and substep of its work item with workgroup barriers only.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "box2d/math_functions.h"

namespace b2mi
{

enum ShapeType : int
{
	kCircle = 0,
	kCapsule = 1,
	kSegment = 2,
	kPolygon = 3,
};

// body reference stored in a shape: awake sim index, or static sim index | kStaticRef
constexpr int kStaticRef = 0x40000000;

// DevBodySim::flags
enum : uint32_t
{
	kBodyDynamic = 1u << 0,
	kBodyKinematic = 1u << 1,
	kBodyBullet = 1u << 2,
	kBodyEnableSleep = 1u << 3,
};

// DevShape::flags
enum : uint32_t
{
	kShapeHitEvents = 1u << 0,
};

// contact sim flags (same values as the reference's src/contact.h:96-115)
enum : uint32_t
{
	kSimTouching = 0x00010000,
	kSimDisjoint = 0x00020000,
	kSimStarted = 0x00040000,
	kSimStopped = 0x00080000,
	kSimHitEvent = 0x00100000,
};

// DevWorld::flags
enum : uint32_t
{
	kWorldWarmStart = 1u << 0,
	kWorldContinuous = 1u << 1,
	kWorldSleep = 1u << 2,
};

// A manifold point as the kernels keep it: the reference's b2ManifoldPoint without the world
// point and the normal velocity, which are output only and live in the contact's second line.
struct DevManifoldPoint
{
	float anchorAX, anchorAY;
	float anchorBX, anchorBY;
	float separation;
	float normalImpulse;
	float tangentImpulse;
	float maxNormalImpulse;
	uint16_t id;
	uint8_t persisted;
	uint8_t pad;
};

struct DevManifold
{
	DevManifoldPoint points[2];
	float normalX, normalY;
	int pointCount;
};

// One 128-byte line: everything the narrowphase reads and writes and the solver prepares from
// (manifold, shapes, flags, material, masses), so each of those kernels touches one line per
// contact. The output-only fields are in DevContactOut.
struct DevContact
{
	DevManifold manifold;
	int shapeIdA, shapeIdB;
	uint32_t simFlags;
	float friction, restitution;
	float invMassA, invIA, invMassB, invIB;
	// For each new manifold point (2 bits each), the index of the previous manifold's point whose
	// impulses it inherited as warm start, or 3 for none; the adjoint routes impulse gradients
	// back through it.
	uint32_t warmFrom;
	int pad;
};
static_assert( sizeof( DevManifoldPoint ) == 36, "manifold point layout" );
static_assert( sizeof( DevContact ) == 128, "one line per contact" );

// The output-only part of a contact, in an array of its own (32 bytes per contact, so the
// narrowphase's writes of neighbouring contacts fill whole lines): the world points and normal
// velocities of the manifold points and the body sim indices.
struct DevContactOut
{
	float pointX[2], pointY[2];
	float normalVelocity[2];
	int bodySimIndexA, bodySimIndexB;
};
static_assert( sizeof( DevContactOut ) == 32, "contact output layout" );

// Two 128-byte lines: the first holds all a shape of up to four vertices gives the narrowphase
// (vertices with their edge normals, type, body, count, radius, fat AABB), so a box pair
// touches one line per shape; vertices 4..7 and the centroid are in the second.
struct DevShape
{
	float vn[4][4]; // vertices 0..3: x, y, edge normal x, y
	int type;
	int body;
	int count;
	float radius;
	float fatAABB[4]; // lower x, lower y, upper x, upper y
	float aabb[4];	  // same order
	uint32_t flags;
	float friction, restitution, pad;
	float vnHigh[4][4]; // vertices 4..7
	float centroid[2];
	float pad2[14];
	// vertex i: [0], [1] position, [2], [3] edge normal
	__host__ __device__ inline float* vert( int i )
	{
		return i < 4 ? vn[i] : vnHigh[i - 4];
	}
	__host__ __device__ inline const float* vert( int i ) const
	{
		return i < 4 ? vn[i] : vnHigh[i - 4];
	}
};
static_assert( sizeof( DevShape ) == 256, "two cache lines" );

// One 128-byte line; the first 32 bytes are what the narrowphase reads.
struct DevBodySim
{
	float p[2], q[2]; // transform: origin and rotation (c, s)
	float localCenter[2];
	float invMass, invI;
	float center[2], center0[2];
	float rotation0[2];
	float force[2];
	float torque, mass;
	float minExtent, maxExtent;
	float linearDamping, angularDamping, gravityScale, sleepThreshold;
	uint32_t flags;
	int shapeStart, shapeCount, pad;
	float sleepTime; // seconds the body has been below its sleep threshold (src/solver.c:245, 277)
	float pad2[3];
};
static_assert( sizeof( DevBodySim ) == 128, "one cache line" );

// The part of a body sim the narrowphase reads (the first 32 bytes of DevBodySim), kept again in
// a compact array of its own so that the narrowphase's gathers touch 32 bytes per body rather
// than a 128-byte line; written where the sims are (upload, finalize).
struct DevPose
{
	float p[2], q[2];
	float localCenter[2];
	float invMass, invI;
};
static_assert( sizeof( DevPose ) == 32, "pose layout" );

// Body state (the reference's 32-byte b2BodyState): two float4 per body
//   s0 = (v.x, v.y, w, flags), s1 = (dp.x, dp.y, dq.c, dq.s)

// ---- joints (src/joint.h: b2JointSim and the per-type structs) ------------------------------

// DevJoint::flags
enum : uint32_t
{
	kJointSpring = 1u << 0,
	kJointLimit = 1u << 1,
	kJointMotor = 1u << 2,
};

struct DevSoft
{
	float biasRate, massScale, impulseScale;
};

// The per-type parts keep the reference's fields (src/joint.h:13-178) except the body indices,
// which a joint's solver slot carries, and the enable flags, which are bits of DevJoint::flags.
struct DevDistanceJoint
{
	float length, hertz, dampingRatio, minLength, maxLength, maxMotorForce, motorSpeed;
	float impulse, lowerImpulse, upperImpulse, motorImpulse;
	b2Vec2 anchorA, anchorB, deltaCenter;
	DevSoft distanceSoftness;
	float axialMass;
};

struct DevMotorJoint
{
	b2Vec2 linearOffset;
	float angularOffset;
	b2Vec2 linearImpulse;
	float angularImpulse;
	float maxForce, maxTorque, correctionFactor;
	b2Vec2 anchorA, anchorB, deltaCenter;
	float deltaAngle;
	b2Mat22 linearMass;
	float angularMass;
};

struct DevMouseJoint
{
	b2Vec2 targetA;
	float hertz, dampingRatio, maxForce;
	b2Vec2 linearImpulse;
	float angularImpulse;
	DevSoft linearSoftness, angularSoftness;
	b2Vec2 anchorB, deltaCenter;
	b2Mat22 linearMass;
};

struct DevPrismaticJoint
{
	b2Vec2 localAxisA, impulse;
	float springImpulse, motorImpulse, lowerImpulse, upperImpulse;
	float hertz, dampingRatio, maxMotorForce, motorSpeed, referenceAngle, lowerTranslation, upperTranslation;
	b2Vec2 anchorA, anchorB, axisA, deltaCenter;
	float deltaAngle, axialMass;
	DevSoft springSoftness;
};

struct DevRevoluteJoint
{
	b2Vec2 linearImpulse;
	float springImpulse, motorImpulse, lowerImpulse, upperImpulse;
	float hertz, dampingRatio, maxMotorTorque, motorSpeed, referenceAngle, lowerAngle, upperAngle;
	b2Vec2 anchorA, anchorB, deltaCenter;
	float deltaAngle, axialMass;
	DevSoft springSoftness;
};

struct DevWeldJoint
{
	float referenceAngle, linearHertz, linearDampingRatio, angularHertz, angularDampingRatio;
	DevSoft linearSoftness, angularSoftness;
	b2Vec2 linearImpulse;
	float angularImpulse;
	b2Vec2 anchorA, anchorB, deltaCenter;
	float deltaAngle, axialMass;
};

struct DevWheelJoint
{
	b2Vec2 localAxisA;
	float perpImpulse, motorImpulse, springImpulse, lowerImpulse, upperImpulse;
	float maxMotorTorque, motorSpeed, lowerTranslation, upperTranslation, hertz, dampingRatio;
	b2Vec2 anchorA, anchorB, axisA, deltaCenter;
	float perpMass, motorMass, axialMass;
	DevSoft springSoftness;
};

// One joint as the solver keeps it, by joint id: parameters, accumulated impulses (warm start
// across steps) and what prepare computes at the start of each step (anchors, masses,
// softness). The solver reads and writes it in place; a joint belongs to one work item.
struct DevJoint
{
	int type; // b2JointType
	uint32_t flags;
	int pad[2];
	b2Vec2 localOriginAnchorA, localOriginAnchorB;
	float invMassA, invMassB, invIA, invIB;
	union
	{
		DevDistanceJoint distance;
		DevMotorJoint motor;
		DevMouseJoint mouse;
		DevPrismaticJoint prismatic;
		DevRevoluteJoint revolute;
		DevWeldJoint weld;
		DevWheelJoint wheel;
		float words[36];
	};
};
static_assert( sizeof( DevJoint ) == 192, "joint record layout" );

struct DevWorld
{
	float* state;		 // [2 * 4 * awakeCapacity]
	DevBodySim* sims;	 // awake set
	DevBodySim* statics; // static set
	DevPose* poses;		 // awake set, the narrowphase's view of sims
	DevPose* staticPoses; // static set, the narrowphase's view of statics
	DevShape* shapes;	 // by shape id
	int* bodyShapes;	 // per awake sim: shape ids, head-first body list order
	DevContact* contacts; // by contact id
	DevContactOut* contactOut; // by contact id
	DevJoint* joints;		   // by joint id
	float gravityX, gravityY;
	float dt, inv_dt, h, inv_h;
	float contactSoft[3]; // biasRate, massScale, impulseScale
	float staticSoft[3];
	float jointSoft[3]; // src/world.c:739-743
	float restitutionThreshold, pushout;
	float maxLinearSpeed, maxAngularSpeed;
	float warmStartScale;
	int subSteps;
	uint32_t flags;
	int pad;
};

// groups of a work item in execution order: 0 = overflow (serial), 1..11 = colours 0..10
constexpr int kGroups = 12;

struct WorkItem
{
	int world;
	int bodyStart;	// into the launch's body list / private state
	int bodyCount;	// owned + shadow (kinematic) bodies
	int ownedCount; // the first ownedCount bodies are finalized and written back
	int slotStart;	// into the launch's constraint slots
	int groupStart[kGroups + 1]; // relative to slotStart
	int ldsBytes;				 // LDS bytes of the item's store (0 for kStoreHbm)
	int store;					 // kStoreHbm, kStoreBodiesLds or kStoreLds
	int block;					 // threads of its workgroup when the launch shape is per item
	// The overflow group in levels: slots of one level share no body whose state they change,
	// so a level runs in parallel and the levels in turn give the serial order's results.
	// Level boundaries (relative to slotStart) at ovfLevels[ovfLevelBegin .. + ovfLevelCount].
	int ovfLevelBegin;
	int ovfLevelCount;
	// For each colour group g (1 .. kGroups-1), the next non-empty colour group after it in the
	// sweep order (cyclically, g itself last), 0 if none: 4 bits at bit 4 (g % 8) of
	// nextGroup[g / 8]. A sweep loads a lane's slot of that group one phase ahead.
	int nextGroup[2];
	// The item's joints follow its contact slots in the slot lists, grouped the same way: slots
	// jointGroupStart[g] .. jointGroupStart[g + 1] (relative to slotStart; jointGroupStart[0] =
	// groupStart[kGroups]) hold the joint ids of group g (0 = overflow, serial in the reference's
	// overflow order, in levels at ovfLevels[jointLevelBegin .. + jointLevelCount]).
	int jointGroupStart[kGroups + 1];
	int jointLevelBegin;
	int jointLevelCount;
};

// every slot of a work item, contacts and joints
__host__ __device__ inline int item_slots( const WorkItem& wi )
{
	return wi.jointGroupStart[kGroups];
}

// the next non-empty colour group after g (see WorkItem::nextGroup)
__host__ __device__ inline int next_group( const WorkItem& wi, int g )
{
	return ( wi.nextGroup[g >> 3] >> ( 4 * ( g & 7 ) ) ) & 15;
}

// The phases of a very large HBM-store work item when it runs as one grid launch per phase
// (k_split_phase) rather than one workgroup: solve_item's stages, in its order.
enum SplitPhase : int
{
	kSplitStage,
	kSplitPrepare,
	kSplitIntegrateVelocities,
	kSplitOverflowWarmStart,
	kSplitWarmStart,
	kSplitOverflowSolve,
	kSplitSolve,
	kSplitIntegratePositions,
	kSplitOverflowRelax,
	kSplitRelax,
	kSplitOverflowRestitution,
	kSplitRestitution,
	kSplitStoreImpulses,
	kSplitFinalize,
	// the joints' stages (one launch per overflow level or colour, before that stage's contacts)
	kSplitJointPrepare,
	kSplitJointWarmStart,
	kSplitJointSolve,
	kSplitJointRelax,
	kSplitPhaseCount
};
// default slot count from which an HBM-store work item runs over the grid (a cooperative launch,
// or one launch per phase where cooperative launches are not available): in a batch, and in a
// group of one world (measured on the 5,050-box pyramid's 14,950-slot island: 1.11 ms of solver
// time in one workgroup, 0.41 ms over the grid)
constexpr int kDefaultSplitSlots = 16384;
constexpr int kDefaultSplitSlotsWorld = 4096;
// threads per workgroup of the cooperative grid solver
constexpr int kGridSolveBlock = 256;
// Overflow levels of an HBM-store item run in runs: consecutive levels of at most
// kLevelCacheSlots slots in all run in one workgroup (of kSplitLevelBlock threads in a grid-wide
// item) with the bodies they touch and their records in LDS, a workgroup barrier per level; a
// larger level runs on its own over all the item's threads.
constexpr int kLevelCacheSlots = 128;
constexpr int kSplitLevelBlock = 256;

// Where a work item keeps its solver state during the step
enum ItemStore : int
{
	kStoreHbm = 0,		 // bodies and constraints in HBM
	kStoreBodiesLds = 1, // bodies in LDS, constraints in HBM
	kStoreLds = 2,		 // bodies and constraints in LDS
};

// narrowphase state change (sorted by world, contactId on the host)
struct ChangeRecord
{
	int world;
	int contactId;
	uint32_t flags;
	int pad;
};

// Broadphase work of finalize and of the continuous stage (sorted by world, sim index on the
// host): a shape whose fat AABB grew (kEnlarged, from finalize; kFastEnlarged, a fast body's
// shape after its sweep), or a fast body (kFastBody, shapeId -1, with kBulletFast for a bullet),
// whose shapes finalize leaves to the continuous stage.
enum : uint32_t
{
	kEnlarged = 1u << 0,
	kFastBody = 1u << 1,
	kBulletFast = 1u << 2,
	kFastEnlarged = 1u << 3,
};

struct EnlargeRecord
{
	int world;
	int simIndex;
	int shapeId;
	uint32_t flags;
	float fatAABB[4];
};


// new contact: device record to scatter into DevWorld::contacts[contactId]
struct ContactInit
{
	int world;
	int contactId;
	int shapeIdA, shapeIdB;
	float friction, restitution;
	int pad[2];
};

// Solver constraint of one slot: a record of seven float4 (112 bytes), so that a lane reads a
// quad of related fields with one 128-bit load and every field sits at a fixed offset:
//   q0  body A, body B (local indices as int bits), normal x, normal y
//   q1  point 0 anchors: A x, A y, B x, B y            q2  the same for point 1
//   q3  point 0 base separation, normal mass, tangent mass; friction
//   q4  point 1 base separation, normal mass, tangent mass; restitution
//   q5  point 0 normal impulse, tangent impulse, max normal impulse, relative velocity
//   q6  the same for point 1
// q0-q4 are read-only during the substep passes; q5 and q6 are written by them.
enum ConstraintField : int
{
	cfBodyA = 0,
	cfBodyB,
	cfNormalX,
	cfNormalY,
	cfAnchorAX = 4, // per point (point 1 at +4)
	cfAnchorAY,
	cfAnchorBX,
	cfAnchorBY,
	cfBaseSeparation = 12, // per point
	cfNormalMass,
	cfTangentMass,
	cfFriction = 15,
	cfRestitution = 19,
	cfNormalImpulse = 20, // per point
	cfTangentImpulse,
	cfMaxNormalImpulse,
	cfRelativeVelocity,
	cfFloatCount = 28,
	cfQuads = 7,
	cfMutableQuad = 5, // first written quad (q5, q6)
	cfMutableCount = 8,
};

// field of manifold point p (0 or 1) for a per-point field given at point 0
constexpr int pointField( int field, int p )
{
	return field + 4 * p;
}

// ---- broadphase pair queries on the device -------------------------------------------------

// A node of a dynamic tree as the pair queries read it (uploaded from the host's trees, which keep
// the reference's node layout and structure): its fat AABB, children, the shape id of a leaf and
// its category bits. A leaf has child1 = -1 and child2 = its position among the tree's leaves in the
// order a depth-first query meets them (child2's subtree before child1's).
struct DevTreeNode
{
	float lowerX, lowerY, upperX, upperY;
	int child1, child2;
	int userData;
	uint32_t categoryBits;
};
static_assert( sizeof( DevTreeNode ) == 32, "tree node layout" );

// What the pair filters read of a shape: its body, the body's type and the shape's filter.
struct DevShapeFilter
{
	int bodyId;
	int bodyType; // b2BodyType
	uint32_t categoryBits, maskBits;
	int groupIndex;
	int pad;
};

// One fast body of the continuous stage (k_continuous) and what its query needs: the trees of
// its world (static, kinematic, dynamic; the latter two for a bullet), the shape filters of its
// world and the bodies its joints keep it from colliding with, noCollide[noCollideBegin ..
// noCollideEnd) in DevShape::body form (sim index, or static sim index | kStaticRef).
struct ContinuousBody
{
	int world;
	int simIndex;
	int noCollideBegin, noCollideEnd;
	const DevTreeNode* nodes[3];
	int roots[3];
	const DevShapeFilter* filters;
};

// the device side of one world's broadphase, passed to k_find_pairs by value
struct DevBroadphase
{
	const DevTreeNode* nodes[3]; // per tree (b2BodyType order)
	int roots[3];
	const DevShapeFilter* shapes; // by shape id
	const uint32_t* moved;		  // bitmap over proxy keys of the moved proxies
	const unsigned long long* pairs; // open-addressing set of the shape pair keys with a contact
	uint32_t pairMask;				 // table size - 1 (a power of two)
};

// A new pair found by the query of move `query`. `rank` orders the pairs of one query as the
// reference's depth-first query meets them: (tree's place in the query order) << 28 | leaf position.
struct PairRecord
{
	int query, rank, shapeIdA, shapeIdB;
};

// A subtree that a query which ran out of its per-thread visit budget had still to visit:
// `entry` = (tree << 30) | node. Each goes on in a workgroup of k_find_pairs_spilled.
struct PairSpill
{
	int query;
	uint32_t entry;
};

// pair-set slots: never a key (shape ids are below 2^31)
constexpr unsigned long long kPairEmpty = ~0ull;
constexpr unsigned long long kPairTombstone = ~0ull - 1;
// threads per workgroup of k_find_pairs, depth of each thread's traversal stack (in LDS) and the
// nodes one thread visits before it spills the rest of its query
constexpr int kPairBlock = 128;
constexpr int kPairStack = 64;
constexpr int kPairBudget = 512;
// k_find_pairs_spilled: threads per workgroup, shared traversal stack entries, workgroups
constexpr int kSpillBlock = 256;
constexpr int kSpillStack = 12288;
constexpr int kSpillGrid = 512;
// the broadphase counters: pairs found, error, spilled subtrees
constexpr int kPairCounters = 3;

// Pose of a body at a step boundary (taped runs): center of mass, rotation, velocities
struct StepPose
{
	float cx, cy, qc, qs;
	float vx, vy, w, pad;
};

// Bytes of LDS a work item's solver state needs: per body two float4 of state and a float2 of
// inverse mass and inertia, per slot its constraint record (which holds its two local bodies).
// Items above the budget keep the read-only constraint fields in HBM, or everything. A CDNA4 CU has
// 160 KiB of LDS, of which the solver kernels' own words (the restitution vote) take the last
// 256 bytes; the default budget leaves room for two or more items per CU.
constexpr int kMaxItemLdsBytes = 160 * 1024 - 256;
constexpr int kDefaultItemLdsBytes = kMaxItemLdsBytes;
// threads per solver workgroup: at most this many (the step kernel's launch bound; the replay
// and adjoint kernels, which hold more state per thread, take at most kAdjointMaxBlock)
constexpr int kSolveMaxBlock = 1024;
constexpr int kAdjointMaxBlock = 256;
// the adjoint kernel keeps a work item's body and slot adjoints (12 floats per body, 8 per slot)
// in LDS when they fit in this many floats, and in its HBM scratch otherwise
constexpr int kAdjointLdsFloats = 4096;
// and its solver state (40 (bodyCount + 1) + 16 cfQuads slots bytes) when it fits in this many bytes
constexpr int kAdjointStoreLdsBytes = 48 * 1024;
// (the LDS stores hold one more body, the static body's identity state, at index bodyCount)
inline int itemLdsBytes( int bodyCount, int slotCount )
{
	return 40 * ( bodyCount + 1 ) + 4 * cfFloatCount * slotCount;
}
// the bodies-in-LDS store: body state and masses, the written quads and the slot bodies
inline int itemBodyLdsBytes( int bodyCount, int slotCount )
{
	return 40 * ( bodyCount + 1 ) + 8 * slotCount + 4 * cfMutableCount * slotCount;
}

} // namespace b2mi
