// Extensions of the MI355X step pipeline beyond the reference's public API.
//
// None of these has a counterpart in the reference (include/box2d/box2d.h); they are
// additions for three kinds of caller:
//  - inspection: tests and tools that compare contact/colour/manifold state with another
//    implementation of the same API (the dump records below are produced identically by
//    this library, by the CPU oracle and by the reference-side tools);
//  - batching: thousands of independent worlds advanced by one device launch per stage;
//  - differentiation: a device adjoint of the soft-step substep behind a ctypes binding.
#pragma once

#include "box2d/box2d.h"

#include <stdint.h>

// ---- inspection --------------------------------------------------------------------------

// One awake contact. Order of records: graph colours 0..11 (11 = overflow) in array order,
// then the awake non-touching set in array order.
typedef struct b2mi_ContactRecord
{
	int32_t contactId;
	int32_t shapeIdA;
	int32_t shapeIdB;
	int32_t colorIndex; // -1 for non-touching
	int32_t localIndex; // index inside its colour array or the non-touching array
	uint32_t simFlags;
	float friction;
	float restitution;
	b2Manifold manifold;
} b2mi_ContactRecord;

// One body, in body-id order (free slots skipped).
typedef struct b2mi_BodyRecord
{
	int32_t bodyId;
	int32_t type;
	int32_t setIndex;	// 0 static, 2 awake
	int32_t localIndex; // index in its solver set
	b2Transform transform;
	b2Vec2 center;
	b2Vec2 linearVelocity;
	float angularVelocity;
	float pad;
} b2mi_BodyRecord;

// Both return the number of records available; at most capacity are written.
B2_API int b2mi_World_DumpContacts( b2WorldId worldId, b2mi_ContactRecord* records, int capacity );
B2_API int b2mi_World_DumpBodies( b2WorldId worldId, b2mi_BodyRecord* records, int capacity );
